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

#include <any>
#include <cstdint>
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

// lib/kascade/configuration.hpp:9-18
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

// lib/kascade/configuration.hpp:39-47.  The MPI-routing knobs are accepted and
// ignored (there is no network on one GPU).
struct PointerDoublingConfig {
  bool use_local_preprocessing = true;
  bool use_grid_communication = false;
  bool use_local_first_request_scheme = true;
  bool cache_succ_owners = true;
  double fallback_allgather_size_ratio = 0.0;
};

enum class RulerSelectionStrategy : std::uint8_t {
  dehne, heuristic, sanders, ultimate, ultimate_sanders, limit_rounds, invalid
};

// lib/kascade/configuration.hpp:65-95.  On the GPU, rulers are one per block of
// `ruler_spacing` ids (0 = the autotuned spacing for the list size) and the
// recursion depth is `sparse_ruling_set_rounds` (0 = until the one-workgroup
// base case).  The ruler-count strategies of the reference size an MPI run and
// are not used.
struct SparseRulingSetConfig {
  RulerSelectionStrategy ruler_selection = RulerSelectionStrategy::sanders;
  Algorithm base_algorithm = Algorithm::PointerDoubling;
  std::any base_algorithm_config;
  std::size_t sparse_ruling_set_rounds = 0;
  std::uint32_t ruler_spacing = 0;
  std::uint32_t ruler_spacing_deep = 0;
  bool validate_list = false;  // the reference's level-100 is_list KASSERT, as a real check
};

// lib/kascade/configuration.hpp:97-103
struct EulerTourConfig {
  Algorithm algorithm = Algorithm::SparseRulingSet;
  bool use_high_degree_handling = false;  // a communication-volume optimisation; ignored
  std::any algo_config;
  bool reproducible_tour = false;  // group children by radix sort instead of atomic linking
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
inline void set_initial_ranking_state(std::span<const idx_t> succ_array, std::span<idx_t> root_array,
                                      std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (root_array.size() != succ_array.size() || rank_array.size() != succ_array.size())
    throw Error(KC_EINVAL, "array sizes differ");
  ctx.check(kc_set_initial_ranking_state(ctx.handle(), succ_array.data(), succ_array.size(),
                                         root_array.data(), rank_array.data(), nullptr));
}

// kascade::pointer_doubling (pointer_doubling.hpp:18-23): in place,
// succ -> root, rank (distance to successor, 0 on roots) -> distance to root.
inline void pointer_doubling(PointerDoublingConfig /*config*/, std::span<idx_t> succ_array,
                             std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (rank_array.size() != succ_array.size()) throw Error(KC_EINVAL, "array sizes differ");
  kc_rank_cfg c = detail::rank_cfg(KC_ALGO_POINTER_DOUBLING);
  ctx.check(kc_pointer_doubling(ctx.handle(), &c, succ_array.data(), rank_array.data(), succ_array.size(),
                                nullptr));
}

// kascade::sparse_ruling_set (sparse_ruling_set.hpp:13-18), in place.
inline void sparse_ruling_set(SparseRulingSetConfig const& config, std::span<idx_t> succ_array,
                              std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (rank_array.size() != succ_array.size()) throw Error(KC_EINVAL, "array sizes differ");
  if (config.base_algorithm == Algorithm::EulerTour || config.base_algorithm == Algorithm::MPLR ||
      config.base_algorithm == Algorithm::invalid)
    throw Error(KC_EINVAL, "Invalid base algorithm selected for sparse ruling set");  // sparse_ruling_set.cpp:374
  if (config.ruler_selection == RulerSelectionStrategy::invalid)
    throw Error(KC_EINVAL, "Invalid ruler selection strategy");  // ruler_selection.hpp:74-75
  kc_rank_cfg c = detail::rank_cfg(KC_ALGO_SPARSE_RULING_SET, config.ruler_spacing, config.ruler_spacing_deep,
                                   static_cast<std::uint32_t>(config.sparse_ruling_set_rounds),
                                   config.validate_list);
  ctx.check(kc_sparse_ruling_set(ctx.handle(), &c, succ_array.data(), rank_array.data(), succ_array.size(),
                                 nullptr));
}

// kascade::rank_via_euler_tour (eulertour.hpp:22-27), in place after
// set_initial_ranking_state: succ = parents -> roots, rank -> depth.
inline void rank_via_euler_tour(EulerTourConfig const& config, std::span<idx_t> succ_array,
                                std::span<rank_t> rank_array, Context& ctx = Context::default_context()) {
  if (rank_array.size() != succ_array.size()) throw Error(KC_EINVAL, "array sizes differ");
  if (config.algorithm == Algorithm::EulerTour) throw Error(KC_EINVAL, "Cannot apply Euler Tour recursively.");
  kc_tree_cfg t{};
  t.rank = detail::rank_cfg(detail::algorithm_code(config.algorithm));
  t.tour_build = config.reproducible_tour ? KC_TOUR_SORTED : KC_TOUR_LINKED;
  ctx.check(kc_rank_via_euler_tour(ctx.handle(), &t, succ_array.data(), rank_array.data(), succ_array.size(),
                                   nullptr));
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

// kascade::compute_euler_tour (eulertour.hpp:32-34): the tour successor array
// in this library's numbering (element 2v = upward edge of v or a root's
// dummy, element 2v + 1 = downward edge into v).
inline auto compute_euler_tour(std::span<const idx_t> parent_array, Context& ctx = Context::default_context())
    -> std::vector<idx_t> {
  std::vector<idx_t> tour(2 * parent_array.size());
  ctx.check(kc_compute_euler_tour(ctx.handle(), parent_array.data(), parent_array.size(), tour.data(), nullptr,
                                  nullptr));
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
    sparse_ruling_set(config_, root_array_, rank_array_, *ctx_);
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
