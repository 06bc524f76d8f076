// kascade_amd.hpp -- C++ drop-in layer over the C ABI (kascade_amd.h), shaped
// like KAscade's lib/kascade entry points and its benchmark plugin interface.
//
// Differences from the reference signatures: there is no kamping::Communicator,
// Distribution or TopologyAwareGridCommunicator argument -- one call ranks the
// whole array on one GPU (the Distribution degenerates to one block).  Code
// written against the reference's in-place contract ports by dropping those
// arguments; `namespace kascade = kascade_amd;` keeps the call sites' names.
//
// Errors: a non-OK kc_status throws kascade_amd::Error (a std::runtime_error,
// like the reference's invalid-configuration throws, e.g.
// pointer_doubling_generic.hpp:446-447); invariant violations that the
// reference only KASSERTs (ids out of range, nonzero root weight, in-degree > 1,
// cycles) are reported the same way instead of being undefined.
#pragma once

#include <algorithm>
#include <any>
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "kascade_amd.h"

namespace kascade_amd {

using idx_t = std::uint64_t;   // kascade::idx_t  (lib/kascade/types.hpp:8)
using rank_t = std::int64_t;   // kascade::rank_t (lib/kascade/types.hpp:9)

namespace bits {  // lib/kascade/bits.hpp:9-21
constexpr idx_t root_flag_mask = idx_t(1) << 63;
[[nodiscard]] constexpr auto set_root_flag(idx_t v) noexcept -> idx_t { return v | root_flag_mask; }
[[nodiscard]] constexpr auto clear_root_flag(idx_t v) noexcept -> idx_t { return v & ~root_flag_mask; }
[[nodiscard]] constexpr auto has_root_flag(idx_t v) noexcept -> bool { return (v & root_flag_mask) != 0; }
}  // namespace bits

// ---- configuration (lib/kascade/configuration.hpp) --------------------------
// Every enum, struct, member and default of the reference, so that code which
// fills these structs compiles unchanged.  Members that only shape MPI
// communication (grid communicator, aggregation, owner caching, request
// schemes, the Briefkasten queue, ruler-count strategies sized for p ranks)
// are accepted and have no effect on one GPU; what each member does here is
// stated beside it.  Members marked "GPU:" are this library's additions.

// configuration.hpp:9-18
enum class Algorithm : std::uint8_t {
  GatherChase,
  PointerDoubling,
  AsyncPointerDoubling,
  RMAPointerDoubling,
  SparseRulingSet,
  EulerTour,
  MPLR,
  invalid,
};

enum class AggregationLevel : std::uint8_t { none, local, all, invalid };  // :20

struct AsyncPointerChasingConfig {  // :22-24 (async doubling = sync doubling here)
  bool use_caching = false;
};

enum class RMASyncMode : std::uint8_t { passive_target, fenced, invalid };  // :26

struct RMAPointerChasingConfig {  // :28-31 (RMA doubling = sync doubling here)
  RMASyncMode sync_mode = RMASyncMode::passive_target;
  std::size_t batch_size = 1;
};

enum class GridCommunicatorMode : std::uint8_t { topology_aware, balanced, invalid };  // :33-37

// :39-47.  use_local_preprocessing = false switches local contraction off
// (on, the library contracts runs of consecutive successors when a sample of
// the input finds them); every other member is an MPI-routing knob:
// accepted, no effect.
struct PointerDoublingConfig {
  bool use_local_preprocessing = true;
  GridCommunicatorMode grid_communicator_mode = GridCommunicatorMode::topology_aware;
  bool use_grid_communication = false;
  bool use_local_first_request_scheme = true;
  bool cache_succ_owners = true;
  AggregationLevel aggregation_level = AggregationLevel::none;
  double fallback_allgather_size_ratio = 0.0;
};

// :49-57
enum class RulerSelectionStrategy : std::uint8_t {
  dehne, heuristic, sanders, ultimate, ultimate_sanders, limit_rounds, invalid
};
enum class RulerPropagationMode : std::uint8_t { pull, push, invalid };  // :58

struct BriefkastenConfig {  // :60-63 (async message queue; no effect)
  std::size_t local_threshold = 2048;
  std::size_t poll_skip_threshold = 1000;
};

// :65-95.
//   sparse_ruling_set_rounds: ruling-set levels before the base algorithm
//     ranks the ruler sublist, as in sparse_ruling_set.cpp:381-392; 0 runs
//     the base algorithm alone (:376-380).
//   base_algorithm: PointerDoubling, AsyncPointerDoubling and
//     RMAPointerDoubling run sync pointer doubling; GatherChase (a gather to
//     one rank and a sequential walk) runs the ruling set's own recursion,
//     whose results are identical; other values throw, as :374 does.
//   ruler_selection and the *_factor members size the ruler set for p MPI
//     ranks; on the GPU the ruler spacing comes from the autotuned table
//     (or ruler_spacing below), so they have no effect, and neither do
//     current_sparse_ruling_set_round, round_limit, ruler_propagation_mode
//     (always pull, the reference's default), sync / sync_locality_aware /
//     no_precompute_rulers / post_invert(_detect_leaves) (the chase engine
//     is this library's own), cache_owners, root_gather_threshold, the
//     aggregation and locality switches, the Briefkasten and grid settings.
//   spawn and use_local_contraction set true force the GPU's counterparts on
//     (ruler spawn in the chase; contraction of runs of consecutive
//     successors); false leaves both to the library.  Results are identical
//     either way.
struct SparseRulingSetConfig {
  RulerSelectionStrategy ruler_selection = RulerSelectionStrategy::sanders;
  Algorithm base_algorithm = Algorithm::PointerDoubling;
  std::any base_algorithm_config;
  std::size_t current_sparse_ruling_set_round = 0;
  std::size_t sparse_ruling_set_rounds = 1;
  RulerPropagationMode ruler_propagation_mode = RulerPropagationMode::pull;
  bool use_aggregation_in_ruler_propagation = false;
  bool use_local_contraction = false;
  bool reverse_list_locality_aware = true;
  bool use_locality_aware_in_ruler_propagation = true;
  double dehne_factor = 1.0;
  double heuristic_factor = 0.01;
  double sanders_factor = 1.0;
  double ultimate_factor = 0.1;
  double ultimate_sanders_factor = 2.0;
  std::size_t round_limit = 100;
  bool cache_owners = false;
  std::size_t root_gather_threshold = 100;
  bool sync = true;
  bool sync_locality_aware = false;
  bool spawn = false;
  bool no_precompute_rulers = false;
  bool post_invert = false;
  bool post_invert_detect_leaves = false;
  BriefkastenConfig briefkasten{};
  GridCommunicatorMode grid_communicator_mode = GridCommunicatorMode::topology_aware;
  bool use_grid_communication = false;
  bool use_local_first_request_scheme = false;
  // GPU: elements per ruler at the top level and below it (powers of two;
  // 0 = the autotuned spacing for the size)
  std::uint32_t ruler_spacing = 0;
  std::uint32_t ruler_spacing_deep = 0;
  // GPU: reject in-degree > 1 before ranking (the reference's is_list KASSERT)
  bool validate_list = false;
};

// :97-103.  algorithm ranks the tour (eulertour.cpp:282-321); algo_config
// may hold the SparseRulingSetConfig whose levels it uses.
// use_high_degree_handling splits high-degree vertices to balance MPI
// messages: no effect, like the grid settings.
struct EulerTourConfig {
  Algorithm algorithm = Algorithm::SparseRulingSet;
  bool use_high_degree_handling = false;
  std::any algo_config;
  GridCommunicatorMode grid_communicator_mode = GridCommunicatorMode::topology_aware;
  bool use_grid_communication = false;
  // GPU: group children by radix sort (the same tour on every call) instead
  // of atomic linking; the results are identical either way
  bool reproducible_tour = false;
};

// The reference's block Distribution (lib/kascade/distribution.hpp:16-36)
// on one GPU: a single block holding every element.  Returned by
// set_initial_ranking_state and accepted (with a communicator and a grid
// communicator of any type) by the entry points' reference-shaped overloads.
class Distribution {
 public:
  Distribution() = default;
  explicit Distribution(std::size_t n) : n_(n) {}
  template <typename Comm>
  Distribution(std::size_t n, Comm const& /*comm*/) : n_(n) {}
  auto get_local_size(int /*rank*/ = 0) const -> std::size_t { return n_; }
  auto get_global_size() const -> std::size_t { return n_; }
  auto get_owner(idx_t /*idx*/) const -> int { return 0; }
  auto get_global_idx(std::size_t local, int /*rank*/ = 0) const -> idx_t { return local; }
  auto get_local_idx(idx_t global, int /*rank*/ = 0) const -> std::size_t { return global; }
  auto index_range_begin(int /*rank*/ = 0) const -> idx_t { return 0; }
  auto index_range_end(int /*rank*/ = 0) const -> idx_t { return n_; }

 private:
  std::size_t n_ = 0;
};

class Error : public std::runtime_error {
 public:
  Error(kc_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  kc_status status;
};

// One device context; the free functions below use a per-thread default
// context on device 0 unless one is passed.
class Context {
 public:
  explicit Context(int device = 0, std::uint64_t mem_budget = 0) {
    kc_status s = kc_context_create(device, mem_budget, &h_);
    if (s != KC_OK) throw Error(s, "kc_context_create failed");
  }
  ~Context() { kc_context_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  void* handle() const { return h_; }
  void check(kc_status s) const {
    if (s != KC_OK) throw Error(s, kc_last_error(h_));
  }
  static Context& default_context() {
    thread_local Context ctx(0);
    return ctx;
  }

 private:
  void* h_ = nullptr;
};

namespace detail {
inline kc_rank_cfg rank_cfg(int algorithm, std::uint32_t spacing = 0, std::uint32_t deep = 0,
                            std::uint32_t levels = 0, bool validate = false) {
  kc_rank_cfg c{};
  c.algorithm = algorithm;
  c.ruler_spacing = spacing;
  c.ruler_spacing_deep = deep;
  c.max_levels = levels;
  c.validate = validate ? 1u : 0u;
  return c;
}
// sparse_ruling_set_rounds and base_algorithm as sparse_ruling_set.cpp:343-392
// read them (see SparseRulingSetConfig).
inline kc_rank_cfg srs_cfg(SparseRulingSetConfig const& config) {
  switch (config.base_algorithm) {
    case Algorithm::GatherChase:
    case Algorithm::PointerDoubling:
    case Algorithm::AsyncPointerDoubling:
    case Algorithm::RMAPointerDoubling:
      break;
    default:
      throw Error(KC_EINVAL, "Invalid base algorithm selected for sparse ruling set");  // sparse_ruling_set.cpp:374
  }
  if (config.ruler_selection == RulerSelectionStrategy::invalid)
    throw Error(KC_EINVAL, "Invalid ruler selection strategy");  // ruler_selection.hpp:74-75
  const bool gather = config.base_algorithm == Algorithm::GatherChase;
  kc_rank_cfg c{};
  if (config.sparse_ruling_set_rounds == 0) {  // the base algorithm alone
    c = rank_cfg(gather ? KC_ALGO_SPARSE_RULING_SET : KC_ALGO_POINTER_DOUBLING, config.ruler_spacing,
                 config.ruler_spacing_deep, 0, config.validate_list);
  } else {
    const std::size_t levels = std::min<std::size_t>(config.sparse_ruling_set_rounds, 64);
    c = rank_cfg(KC_ALGO_SPARSE_RULING_SET, config.ruler_spacing, config.ruler_spacing_deep,
                 gather ? 0u : static_cast<std::uint32_t>(levels), config.validate_list);
  }
  // Results are identical either way: true forces the GPU's counterpart on,
  // false leaves the choice to the library (a sample of the input decides).
  c.local_contraction = config.use_local_contraction ? KC_CONTRACT_ON : KC_CONTRACT_AUTO;
  c.spawn = config.spawn ? KC_SPAWN_ON : KC_SPAWN_AUTO;
  return c;
}
inline int algorithm_code(Algorithm a) {
  switch (a) {
    case Algorithm::PointerDoubling:
    case Algorithm::AsyncPointerDoubling:  // same outputs as sync doubling
    case Algorithm::RMAPointerDoubling:
      return KC_ALGO_POINTER_DOUBLING;
    case Algorithm::SparseRulingSet:
    case Algorithm::GatherChase:  // outputs are unique; ranked by the ruling set
      return KC_ALGO_SPARSE_RULING_SET;
    default:
      throw Error(KC_EINVAL, "No suitable algorithm selected");  // eulertour.cpp:318
  }
}
}  // namespace detail

// kascade::set_initial_ranking_state (list_ranking.hpp:47-50)
inline auto set_initial_ranking_state(std::span<const idx_t> succ_array, std::span<idx_t> root_array,
                                      std::span<rank_t> rank_array, Context& ctx = Context::default_context())
    -> Distribution {
  if (root_array.size() != succ_array.size() || rank_array.size() != succ_array.size())
    throw Error(KC_EINVAL, "array sizes differ");
  ctx.check(kc_set_initial_ranking_state(ctx.handle(), succ_array.data(), succ_array.size(),
                                         root_array.data(), rank_array.data(), nullptr));
  return Distribution(succ_array.size());
}
// The reference's signature (list_ranking.hpp:47-50): the communicator, of
// any type, is ignored; the default context ranks.
template <typename Comm>
  requires(!std::is_same_v<std::remove_cvref_t<Comm>, Context>)
auto set_initial_ranking_state(std::span<const idx_t> succ_array, std::span<idx_t> root_array,
                               std::span<rank_t> rank_array, Comm const& /*comm*/) -> Distribution {
  return set_initial_ranking_state(succ_array, root_array, rank_array, Context::default_context());
}

// kascade::pointer_doubling (pointer_doubling.hpp:18-23): in place,
// succ -> root, rank (distance to successor, 0 on roots) -> distance to root.
inline void pointer_doubling(PointerDoublingConfig config, std::span<idx_t> succ_array,
                             std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (rank_array.size() != succ_array.size()) throw Error(KC_EINVAL, "array sizes differ");
  kc_rank_cfg c = detail::rank_cfg(KC_ALGO_POINTER_DOUBLING);
  // pointer_doubling_generic.hpp:639: the local preprocessing contracts runs
  // before the doubling; off, nothing is contracted
  if (!config.use_local_preprocessing) c.local_contraction = KC_CONTRACT_OFF;
  ctx.check(kc_pointer_doubling(ctx.handle(), &c, succ_array.data(), rank_array.data(), succ_array.size(),
                                nullptr));
}

// kascade::sparse_ruling_set (sparse_ruling_set.hpp:13-18), in place.
inline void sparse_ruling_set(SparseRulingSetConfig const& config, std::span<idx_t> succ_array,
                              std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (rank_array.size() != succ_array.size()) throw Error(KC_EINVAL, "array sizes differ");
  kc_rank_cfg c = detail::srs_cfg(config);
  ctx.check(kc_rank_list(ctx.handle(), succ_array.data(), rank_array.data(), succ_array.size(), succ_array.data(),
                         rank_array.data(), &c, nullptr));
}

// kascade::rank_via_euler_tour (eulertour.hpp:22-27), in place after
// set_initial_ranking_state: succ = parents -> roots, rank -> depth.
inline void rank_via_euler_tour(EulerTourConfig const& config, std::span<idx_t> succ_array,
                                std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (rank_array.size() != succ_array.size()) throw Error(KC_EINVAL, "array sizes differ");
  if (config.algorithm == Algorithm::EulerTour) throw Error(KC_EINVAL, "Cannot apply Euler Tour recursively.");
  kc_tree_cfg t{};
  t.rank = detail::rank_cfg(detail::algorithm_code(config.algorithm));
  if (config.algorithm == Algorithm::SparseRulingSet)
    if (auto const* srs = std::any_cast<SparseRulingSetConfig>(&config.algo_config)) t.rank = detail::srs_cfg(*srs);
  t.tour_build = config.reproducible_tour ? KC_TOUR_SORTED : KC_TOUR_LINKED;
  ctx.check(kc_rank_via_euler_tour(ctx.handle(), &t, succ_array.data(), rank_array.data(), succ_array.size(),
                                   nullptr));
}

// The reference's six-argument signatures (pointer_doubling.hpp:18-23,
// sparse_ruling_set.hpp:13-18, eulertour.hpp:22-27).  One GPU ranks the
// whole array, so the call must come from a single process holding all of
// it: a communicator with a size() above 1 throws (gather to one rank first,
// as GatherRank does, list_ranking.cpp:29-58).  The Distribution and the
// optional grid communicator, of any type, are otherwise ignored, so call
// sites compile unchanged.
namespace detail {
template <typename Comm>
void require_one_process(Comm const& comm) {
  if constexpr (requires { comm.size(); }) {
    if (comm.size() != 1)
      throw Error(KC_EINVAL, "the whole array must be on one process: gather it to one rank first");
  }
}
}  // namespace detail
template <typename Dist, typename Comm, typename Grid>
void pointer_doubling(PointerDoublingConfig config, std::span<idx_t> succ_array, std::span<rank_t> rank_array,
                      Dist const& /*dist*/, Comm const& comm, Grid const& /*grid_comm*/) {
  detail::require_one_process(comm);
  pointer_doubling(config, succ_array, rank_array);
}
template <typename Dist, typename Comm, typename Grid>
void sparse_ruling_set(SparseRulingSetConfig const& config, std::span<idx_t> succ_array,
                       std::span<rank_t> rank_array, Dist const& /*dist*/, Comm const& comm,
                       Grid const& /*grid_comm*/) {
  detail::require_one_process(comm);
  sparse_ruling_set(config, succ_array, rank_array);
}
template <typename Dist, typename Comm, typename Grid>
void rank_via_euler_tour(EulerTourConfig const& config, std::span<idx_t> succ_array, std::span<rank_t> rank_array,
                         Dist const& /*dist*/, Comm const& comm, Grid const& /*grid_comm*/) {
  detail::require_one_process(comm);
  rank_via_euler_tour(config, succ_array, rank_array);
}

// Conversions from the reference's own configuration types (any types with
// the members of configuration.hpp:39-103; enums convert by value, their
// enumerators being in the same order), so a harness built against both
// libraries passes its kascade::*Config objects straight through.
namespace detail {
template <typename To, typename From>
To enum_cast(From f) {
  return static_cast<To>(static_cast<std::underlying_type_t<From>>(f));
}
}  // namespace detail
template <typename Ref>
  requires requires(Ref const& r) { r.fallback_allgather_size_ratio; r.cache_succ_owners; }
auto from_reference(Ref const& r) -> PointerDoublingConfig {
  PointerDoublingConfig c;
  c.use_local_preprocessing = r.use_local_preprocessing;
  c.grid_communicator_mode = detail::enum_cast<GridCommunicatorMode>(r.grid_communicator_mode);
  c.use_grid_communication = r.use_grid_communication;
  c.use_local_first_request_scheme = r.use_local_first_request_scheme;
  c.cache_succ_owners = r.cache_succ_owners;
  c.aggregation_level = detail::enum_cast<AggregationLevel>(r.aggregation_level);
  c.fallback_allgather_size_ratio = r.fallback_allgather_size_ratio;
  return c;
}
template <typename Ref>
  requires requires(Ref const& r) { r.sparse_ruling_set_rounds; r.ruler_selection; }
auto from_reference(Ref const& r) -> SparseRulingSetConfig {
  SparseRulingSetConfig c;
  c.ruler_selection = detail::enum_cast<RulerSelectionStrategy>(r.ruler_selection);
  c.base_algorithm = detail::enum_cast<Algorithm>(r.base_algorithm);
  c.current_sparse_ruling_set_round = r.current_sparse_ruling_set_round;
  c.sparse_ruling_set_rounds = r.sparse_ruling_set_rounds;
  c.ruler_propagation_mode = detail::enum_cast<RulerPropagationMode>(r.ruler_propagation_mode);
  c.use_aggregation_in_ruler_propagation = r.use_aggregation_in_ruler_propagation;
  c.use_local_contraction = r.use_local_contraction;
  c.reverse_list_locality_aware = r.reverse_list_locality_aware;
  c.use_locality_aware_in_ruler_propagation = r.use_locality_aware_in_ruler_propagation;
  c.dehne_factor = r.dehne_factor;
  c.heuristic_factor = r.heuristic_factor;
  c.sanders_factor = r.sanders_factor;
  c.ultimate_factor = r.ultimate_factor;
  c.ultimate_sanders_factor = r.ultimate_sanders_factor;
  c.round_limit = r.round_limit;
  c.cache_owners = r.cache_owners;
  c.root_gather_threshold = r.root_gather_threshold;
  c.sync = r.sync;
  c.sync_locality_aware = r.sync_locality_aware;
  c.spawn = r.spawn;
  c.no_precompute_rulers = r.no_precompute_rulers;
  c.post_invert = r.post_invert;
  c.post_invert_detect_leaves = r.post_invert_detect_leaves;
  c.briefkasten.local_threshold = r.briefkasten.local_threshold;
  c.briefkasten.poll_skip_threshold = r.briefkasten.poll_skip_threshold;
  c.grid_communicator_mode = detail::enum_cast<GridCommunicatorMode>(r.grid_communicator_mode);
  c.use_grid_communication = r.use_grid_communication;
  c.use_local_first_request_scheme = r.use_local_first_request_scheme;
  return c;
}
// RefSrs: the reference's SparseRulingSetConfig type, which algo_config may hold.
template <typename RefSrs, typename Ref>
  requires requires(Ref const& r) { r.use_high_degree_handling; r.algo_config; }
auto from_reference(Ref const& r) -> EulerTourConfig {
  EulerTourConfig c;
  c.algorithm = detail::enum_cast<Algorithm>(r.algorithm);
  c.use_high_degree_handling = r.use_high_degree_handling;
  if (auto const* srs = std::any_cast<RefSrs>(&r.algo_config)) c.algo_config = from_reference(*srs);
  c.grid_communicator_mode = detail::enum_cast<GridCommunicatorMode>(r.grid_communicator_mode);
  c.use_grid_communication = r.use_grid_communication;
  return c;
}

// kascade::is_list (successor_utils.hpp): no element has two predecessors.
inline auto is_list(std::span<const idx_t> succ_array, Context& ctx = Context::default_context()) -> bool {
  std::int32_t out = 0;
  ctx.check(kc_is_list(ctx.handle(), succ_array.data(), succ_array.size(), &out, nullptr));
  return out != 0;
}

// kascade::LeafInfo (successor_utils.cpp:72-123): is_leaf(v) = no predecessor
// (a root without predecessor counts as a leaf).
class LeafInfo {
 public:
  explicit LeafInfo(std::span<const idx_t> succ_array, Context& ctx = Context::default_context())
      : is_leaf_(succ_array.size()) {
    ctx.check(kc_leaf_info(ctx.handle(), succ_array.data(), succ_array.size(), is_leaf_.data(), &num_leaves_,
                           nullptr));
  }
  auto is_leaf(idx_t idx) const -> bool { return is_leaf_[idx] != 0; }
  auto num_local_leaves() const -> std::size_t { return num_leaves_; }

 private:
  std::vector<std::uint8_t> is_leaf_;
  std::uint64_t num_leaves_ = 0;
};

// kascade::reverse_list (successor_utils.cpp:159-232): fills pred_array and
// dist_to_pred; returns the original roots (now leaves) in ascending order.
inline auto reverse_list(std::span<const idx_t> succ_array, std::span<const rank_t> dist_to_succ,
                         std::span<idx_t> pred_array, std::span<rank_t> dist_to_pred,
                         Context& ctx = Context::default_context()) -> std::vector<idx_t> {
  const std::size_t n = succ_array.size();
  if (pred_array.size() < n || dist_to_pred.size() < n || (!dist_to_succ.empty() && dist_to_succ.size() < n))
    throw Error(KC_EINVAL, "array sizes do not match");
  std::vector<idx_t> roots(n);
  std::uint64_t nroots = 0;
  ctx.check(kc_reverse_list(ctx.handle(), succ_array.data(), dist_to_succ.empty() ? nullptr : dist_to_succ.data(),
                            n, pred_array.data(), dist_to_pred.data(), roots.data(), n, &nroots, nullptr));
  roots.resize(nroots);
  return roots;
}

// compute_stats (bench/detail/stats.cpp:147-162) with the reference's
// StatsLevel and result shapes (bench/detail/stats.hpp:9-44).
enum class StatsLevel : std::uint8_t { invalid, none, basic, reduced_extensive, extensive };
struct BasicStats {
  std::size_t num_trees = 0;
  std::int64_t max_rank = 0;
  double avg_rank = 0.;
};
struct TreeStats {
  idx_t root = 0;
  std::size_t size = 0;
  std::int64_t max_rank = 0;
  std::int64_t rank_sum = 0;
};
struct ExtensiveStats {
  std::size_t num_trees = 0;
  std::size_t num_nontrivial_trees = 0;
  std::size_t nontrivial_size_sum = 0;
  std::int64_t nontrivial_rank_sum = 0;
  std::size_t max_size = 0;
  double avg_size = 0.;
  double nontrivial_avg_size = 0.;
  double nontrivial_avg_rank = 0.;
  std::optional<std::vector<TreeStats>> per_tree_stats;  // ascending root order
};
struct Stats {
  std::optional<BasicStats> basic_stats;
  std::optional<ExtensiveStats> extensive_stats;
};

inline auto compute_stats(StatsLevel level, std::span<const idx_t> root_array, std::span<const rank_t> rank_array,
                          Context& ctx = Context::default_context()) -> Stats {
  if (level == StatsLevel::invalid) throw Error(KC_EINVAL, "invalid statistics level");
  if (level == StatsLevel::none) return Stats{};
  const std::size_t n = root_array.size();
  kc_stats st{};
  std::vector<idx_t> tr;
  std::vector<std::uint64_t> ts;
  std::vector<std::int64_t> tm, tsum;
  if (level == StatsLevel::extensive) {
    tr.resize(n);
    ts.resize(n);
    tm.resize(n);
    tsum.resize(n);
  }
  ctx.check(kc_compute_stats(ctx.handle(), static_cast<std::int32_t>(level), root_array.data(), rank_array.data(),
                             n, &st, tr.data(), ts.data(), tm.data(), tsum.data(), tr.size(), nullptr));
  Stats out;
  out.basic_stats = BasicStats{st.num_trees, st.max_rank, st.avg_rank};
  if (level == StatsLevel::basic) return out;
  ExtensiveStats e{st.num_trees, st.num_nontrivial_trees, st.nontrivial_size_sum, st.nontrivial_rank_sum,
                   st.max_size, st.avg_size, st.nontrivial_avg_size, st.nontrivial_avg_rank, std::nullopt};
  if (level == StatsLevel::extensive) {
    std::vector<TreeStats> per(st.num_nontrivial_trees);
    for (std::size_t i = 0; i < per.size(); i++) per[i] = TreeStats{tr[i], ts[i], tm[i], tsum[i]};
    e.per_tree_stats = std::move(per);
  }
  out.extensive_stats = std::move(e);
  return out;
}

// kascade::compute_euler_tour (eulertour.hpp:32-34, eulertour.cpp:379-392):
// the tour successor array in the reference's element order and length
// (2(n - roots) + roots; see kc_compute_euler_tour).
inline auto compute_euler_tour(std::span<const idx_t> parent_array, Context& ctx = Context::default_context())
    -> std::vector<idx_t> {
  std::vector<idx_t> tour(2 * parent_array.size());
  std::uint64_t len = 0;
  ctx.check(kc_compute_euler_tour(ctx.handle(), parent_array.data(), parent_array.size(), tour.data(), nullptr,
                                  &len, nullptr));
  tour.resize(len);
  return tour;
}

// kascade::pack / unpack (pack.hpp:23-177) on one address space.
inline void pack(std::span<const idx_t> succ_array, std::span<const rank_t> rank_array,
                 std::span<const idx_t> active_indices, std::span<idx_t> succ_array_packed,
                 std::span<rank_t> rank_array_packed, Context& ctx = Context::default_context()) {
  if (succ_array_packed.size() < active_indices.size() || rank_array_packed.size() < active_indices.size())
    throw Error(KC_EINVAL, "packed arrays too small");
  ctx.check(kc_pack(ctx.handle(), succ_array.data(), rank_array.data(), succ_array.size(), active_indices.data(),
                    active_indices.size(), succ_array_packed.data(), rank_array_packed.data(), nullptr));
}
inline void unpack(std::span<const idx_t> succ_array_packed, std::span<const rank_t> rank_array_packed,
                   std::span<const idx_t> active_indices, std::span<idx_t> succ_array,
                   std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  ctx.check(kc_unpack(ctx.handle(), succ_array_packed.data(), rank_array_packed.data(), active_indices.data(),
                      active_indices.size(), succ_array.data(), rank_array.data(), succ_array.size(), nullptr));
}

// ---- the benchmark plugin interface (bench/detail/algorithm.hpp:7-20) ------
class AbstractAlgorithm {
 public:
  virtual ~AbstractAlgorithm() = default;
  virtual void ingest(std::span<const idx_t> succ_array) = 0;
  virtual void run() = 0;
  virtual auto get_rank_array() -> std::vector<rank_t> const& = 0;
  virtual auto get_root_array() -> std::vector<idx_t> const& = 0;
};

// AlgorithmBase (bench/detail/algorithm_impl.hpp:59-99): copies of the input,
// result arrays owned by the adapter; run() starts from
// set_initial_ranking_state like every reference adapter.
class AlgorithmBase : public AbstractAlgorithm {
 public:
  explicit AlgorithmBase(Context& ctx = Context::default_context()) : ctx_(&ctx) {}
  void ingest(std::span<const idx_t> succ_array) override {
    succ_array_ = succ_array;
    rank_array_.resize(succ_array.size());
    root_array_.resize(succ_array.size());
  }
  auto get_rank_array() -> std::vector<rank_t> const& override { return rank_array_; }
  auto get_root_array() -> std::vector<idx_t> const& override { return root_array_; }

 protected:
  void init() { set_initial_ranking_state(succ_array_, root_array_, rank_array_, *ctx_); }
  std::span<const idx_t> succ_array_;
  std::vector<rank_t> rank_array_;
  std::vector<idx_t> root_array_;
  Context* ctx_;
};

class PointerDoubling : public AlgorithmBase {  // algorithm_impl.hpp:111-129
 public:
  explicit PointerDoubling(PointerDoublingConfig c = {}, Context& ctx = Context::default_context())
      : AlgorithmBase(ctx), config_(c) {}
  void run() override {
    init();
    pointer_doubling(config_, root_array_, rank_array_, *ctx_);
  }

 private:
  PointerDoublingConfig config_;
};

class SparseRulingSet : public AlgorithmBase {  // algorithm_impl.hpp:161-201
 public:
  explicit SparseRulingSet(SparseRulingSetConfig c = {}, Context& ctx = Context::default_context())
      : AlgorithmBase(ctx), config_(std::move(c)) {}
  void run() override {
    init();
    SparseRulingSetConfig c = config_;
    if (c.base_algorithm == Algorithm::SparseRulingSet) {
      // the legacy base algorithm (algorithm_impl.hpp:166-171): at least two
      // ruling-set levels over a pointer-doubling base
      c.sparse_ruling_set_rounds = std::max(std::size_t{2}, c.sparse_ruling_set_rounds);
      c.base_algorithm = Algorithm::PointerDoubling;
    }
    sparse_ruling_set(c, root_array_, rank_array_, *ctx_);
  }

 private:
  SparseRulingSetConfig config_;
};

class EulerTour : public AlgorithmBase {  // algorithm_impl.hpp:203-246
 public:
  explicit EulerTour(EulerTourConfig c = {}, Context& ctx = Context::default_context())
      : AlgorithmBase(ctx), config_(std::move(c)) {}
  void run() override {
    init();
    rank_via_euler_tour(config_, root_array_, rank_array_, *ctx_);
  }

 private:
  EulerTourConfig config_;
};

}  // namespace kascade_amd
