// Self-checking C++ test of the drop-in header (include/kascade_amd.hpp),
// written like the reference's tests/pack_test.cpp and its harness adapters:
// the pack known-answer test on one device, the identity sanity check, and
// every AbstractAlgorithm adapter verified element-wise against the oracle's
// sequential walk (the harness's GatherRank reference, bench/detail/
// verification.cpp:9-88).  Exit status 0 = all checks passed.
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <numeric>
#include <optional>
#include <vector>

#include "kascade_amd.hpp"

extern "C" {  // CPU oracle (oracle/oracle.c), the checker
int64_t orc_compute_euler_tour(const uint64_t* parent, uint64_t n, uint64_t* tsucc, int64_t* tw);
void orc_gen_list(uint64_t n, uint64_t seed, int weighted, uint64_t* succ, int64_t* w);
void orc_gen_list_local(uint64_t n, uint64_t seed, int weighted, uint64_t* succ, int64_t* w);
void orc_gen_tree(uint64_t n, uint64_t seed, uint64_t* parent);
void orc_local_pointer_chasing(uint64_t* succ, int64_t* rank, uint64_t n);
int orc_euler_tour_rooting(uint64_t* succ, int64_t* rank, uint64_t n);
int orc_is_list(const uint64_t* succ, uint64_t n);
int64_t orc_reverse_list(const uint64_t* succ, const int64_t* dist, uint64_t n, uint64_t* pred,
                         int64_t* dist_to_pred, uint64_t* roots);
}

namespace ka = kascade_amd;
static int failures = 0;
#define EXPECT(cond)                                                      \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::fprintf(stderr, "%s:%d: EXPECT(%s) failed\n", __FILE__, __LINE__, #cond); \
      failures++;                                                         \
    }                                                                     \
  } while (0)

static void pack_works() {
  // tests/pack_test.cpp:14-109 with the four ranks' arrays laid end to end
  // (offsets 0, 8, 14, 16, 20)
  std::vector<ka::idx_t> succ = {0, 6, 0, 0, 0, 0, 8, 0, 10, 0, 14, 0, 0, 0, 15, 18, 0, 0, 18, 0};
  std::vector<ka::rank_t> rank = {42, 1, 42, 42, 42, 42, 1, 42, 1, 42, 1, 42, 42, 42, 1, 1, 42, 42, 0, 42};
  std::vector<ka::idx_t> active = {1, 6, 8, 10, 14, 15, 18};
  std::vector<ka::idx_t> psucc(7, 42);
  std::vector<ka::rank_t> prank(7, 42);
  ka::pack(succ, rank, active, psucc, prank);
  EXPECT((psucc == std::vector<ka::idx_t>{1, 2, 3, 4, 5, 6, 6}));
  EXPECT((prank == std::vector<ka::rank_t>{1, 1, 1, 1, 1, 1, 0}));
  // simulated ranking: root 6, rank 6 - i - 2r for local packed index i on rank r
  std::fill(psucc.begin(), psucc.end(), 6);
  prank = {6, 5, 4, 3, 2, 1, 0};
  ka::unpack(psucc, prank, active, succ, rank);
  EXPECT((succ == std::vector<ka::idx_t>{0, 18, 0, 0, 0, 0, 18, 0, 18, 0, 18, 0, 0, 0, 18, 18, 0, 0, 18, 0}));
  EXPECT((rank == std::vector<ka::rank_t>{42, 6, 42, 42, 42, 42, 5, 42, 4, 42, 3, 42, 42, 42, 2, 1, 42, 42, 0, 42}));
}

static void pack_sanity_check() {
  // tests/pack_test.cpp:111-135: a 12-element path packed with every index active
  std::vector<ka::idx_t> succ(12);
  std::iota(succ.begin(), succ.end(), 1);
  succ.back() = 11;
  std::vector<ka::rank_t> rank(12, 1);
  rank.back() = 0;
  std::vector<ka::idx_t> active(12);
  std::iota(active.begin(), active.end(), 0);
  std::vector<ka::idx_t> psucc(12, 42);
  std::vector<ka::rank_t> prank(12, 42);
  ka::pack(succ, rank, active, psucc, prank);
  EXPECT(psucc == succ);
  EXPECT(prank == rank);
}

static void adapters_match_gather_rank() {
  for (uint64_t n : {1ull, 2ull, 1000ull, 70001ull}) {
    std::vector<ka::idx_t> succ(n);
    std::vector<ka::rank_t> w(n);
    orc_gen_list(n, 99 + n, 0, succ.data(), w.data());
    std::vector<ka::idx_t> eroot = succ;
    std::vector<ka::rank_t> erank(n);
    for (uint64_t i = 0; i < n; i++) erank[i] = succ[i] == i ? 0 : 1;  // set_initial_ranking_state
    orc_local_pointer_chasing(eroot.data(), erank.data(), n);         // GatherRank
    ka::PointerDoubling pd;
    ka::SparseRulingSet srs;
    for (ka::AbstractAlgorithm* a : {static_cast<ka::AbstractAlgorithm*>(&pd), static_cast<ka::AbstractAlgorithm*>(&srs)}) {
      a->ingest(succ);
      a->run();
      EXPECT(a->get_root_array() == eroot);
      EXPECT(a->get_rank_array() == erank);
    }
    // weighted, in place (pointer_doubling.hpp:14-17 contract)
    orc_gen_list(n, 7 + n, 1, succ.data(), w.data());
    std::vector<ka::idx_t> s2 = succ;
    std::vector<ka::rank_t> r2 = w;
    orc_local_pointer_chasing(s2.data(), r2.data(), n);
    std::vector<ka::idx_t> s3 = succ;
    std::vector<ka::rank_t> r3 = w;
    ka::sparse_ruling_set({}, s3, r3);
    EXPECT(s3 == s2 && r3 == r2);
    s3 = succ;
    r3 = w;
    ka::pointer_doubling({}, s3, r3);
    EXPECT(s3 == s2 && r3 == r2);
  }
  for (uint64_t n : {1ull, 5ull, 30011ull}) {
    std::vector<ka::idx_t> parent(n);
    orc_gen_tree(n, 3 + n, parent.data());
    std::vector<ka::idx_t> eroot = parent;
    std::vector<ka::rank_t> edepth(n);
    orc_euler_tour_rooting(eroot.data(), edepth.data(), n);
    ka::EulerTour et;
    et.ingest(parent);
    et.run();
    EXPECT(et.get_root_array() == eroot);
    EXPECT(et.get_rank_array() == edepth);
  }
}

static void errors_throw() {
  std::vector<ka::idx_t> bad = {1, 5, 2};
  std::vector<ka::rank_t> r = {1, 1, 0};
  bool thrown = false;
  try {
    ka::sparse_ruling_set({}, bad, r);
  } catch (const ka::Error& e) {
    thrown = e.status == KC_EINVAL;
  }
  EXPECT(thrown);
  ka::SparseRulingSetConfig c;
  c.base_algorithm = ka::Algorithm::EulerTour;
  thrown = false;
  try {
    std::vector<ka::idx_t> s = {0};
    std::vector<ka::rank_t> k = {0};
    ka::sparse_ruling_set(c, s, k);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  EXPECT(thrown);
}

// successor utilities and statistics through the C++ layer (successor_utils.hpp,
// bench/detail/stats.hpp shapes)
static void utilities_match_oracle() {
  const uint64_t n = 50000;
  std::vector<ka::idx_t> succ(n);
  std::vector<ka::rank_t> w(n);
  orc_gen_list(n, 31, 1, succ.data(), w.data());
  EXPECT(ka::is_list(succ) == (orc_is_list(succ.data(), n) == 1));
  std::vector<ka::idx_t> two = {2, 2, 2};
  EXPECT(!ka::is_list(two));
  ka::LeafInfo info(succ);
  EXPECT(info.num_local_leaves() == 1);
  std::vector<ka::idx_t> pred(n), epred(n), eroots(n);
  std::vector<ka::rank_t> dpred(n), edpred(n);
  auto roots = ka::reverse_list(succ, w, pred, dpred);
  const int64_t k = orc_reverse_list(succ.data(), w.data(), n, epred.data(), edpred.data(), eroots.data());
  eroots.resize(static_cast<size_t>(k));
  EXPECT(pred == epred);
  EXPECT(dpred == edpred);
  EXPECT(roots == eroots);
  std::vector<ka::idx_t> root = succ;
  std::vector<ka::rank_t> rank = w;
  orc_local_pointer_chasing(root.data(), rank.data(), n);
  auto st = ka::compute_stats(ka::StatsLevel::extensive, root, rank);
  EXPECT(st.basic_stats && st.basic_stats->num_trees == 1);
  EXPECT(st.extensive_stats && st.extensive_stats->per_tree_stats->size() == 1);
  int64_t sum = 0, mx = 0;
  for (auto r : rank) {
    sum += r;
    mx = r > mx ? r : mx;
  }
  EXPECT(st.basic_stats->max_rank == mx);
  EXPECT((*st.extensive_stats->per_tree_stats)[0].rank_sum == sum);
  EXPECT((*st.extensive_stats->per_tree_stats)[0].size == n);
}

// Code written against lib/kascade's configuration and entry points, with
// `namespace kascade = kascade_amd;`: a harness-style adapter fills every
// configuration member the reference declares (configuration.hpp:39-103),
// stores configs in the std::any members as bench/detail/algorithm_impl.hpp
// does, and calls the six-argument entry points with a Distribution, a
// communicator and a grid communicator of its own types.
namespace kascade = kascade_amd;
struct FakeComm {
  int size_ = 1;
  int rank() const { return 0; }
  int size() const { return size_; }
};
struct FakeGrid {};
// A stand-in for the reference's own configuration type (same members).
struct RefSrsConfig : kascade_amd::SparseRulingSetConfig {};

static void reference_shaped_calls() {
  kascade::PointerDoublingConfig pd;
  pd.use_local_preprocessing = true;
  pd.grid_communicator_mode = kascade::GridCommunicatorMode::balanced;
  pd.use_grid_communication = false;
  pd.use_local_first_request_scheme = true;
  pd.cache_succ_owners = true;
  pd.aggregation_level = kascade::AggregationLevel::local;
  pd.fallback_allgather_size_ratio = 3.0;
  kascade::AsyncPointerChasingConfig apc;
  apc.use_caching = true;
  kascade::RMAPointerChasingConfig rma;
  rma.sync_mode = kascade::RMASyncMode::fenced;
  rma.batch_size = 16;
  kascade::SparseRulingSetConfig srs;
  EXPECT(srs.sparse_ruling_set_rounds == 1);  // the reference's default
  srs.ruler_selection = kascade::RulerSelectionStrategy::ultimate_sanders;
  srs.base_algorithm = kascade::Algorithm::PointerDoubling;
  srs.base_algorithm_config = pd;
  srs.base_algorithm_config = apc;
  srs.base_algorithm_config = rma;
  srs.base_algorithm_config = nullptr;
  srs.current_sparse_ruling_set_round = 0;
  srs.sparse_ruling_set_rounds = std::max(std::size_t{2}, srs.sparse_ruling_set_rounds);
  srs.ruler_propagation_mode = kascade::RulerPropagationMode::pull;
  srs.use_aggregation_in_ruler_propagation = false;
  srs.use_local_contraction = true;
  srs.reverse_list_locality_aware = true;
  srs.use_locality_aware_in_ruler_propagation = true;
  srs.dehne_factor = 1.0;
  srs.heuristic_factor = 0.01;
  srs.sanders_factor = 1.0;
  srs.ultimate_factor = 0.1;
  srs.ultimate_sanders_factor = 2.0;
  srs.round_limit = 100;
  srs.cache_owners = true;
  srs.root_gather_threshold = 100;
  srs.sync = true;
  srs.sync_locality_aware = true;
  srs.spawn = true;
  srs.no_precompute_rulers = false;
  srs.post_invert = true;
  srs.post_invert_detect_leaves = false;
  srs.briefkasten.local_threshold = 2048;
  srs.briefkasten.poll_skip_threshold = 1000;
  srs.grid_communicator_mode = kascade::GridCommunicatorMode::topology_aware;
  srs.use_grid_communication = false;
  srs.use_local_first_request_scheme = false;
  kascade::EulerTourConfig et;
  et.algorithm = kascade::Algorithm::SparseRulingSet;
  et.use_high_degree_handling = false;
  et.algo_config = srs;
  et.grid_communicator_mode = kascade::GridCommunicatorMode::topology_aware;
  et.use_grid_communication = false;

  FakeComm comm;
  std::optional<FakeGrid> grid;
  const uint64_t n = 60013;
  std::vector<kascade::idx_t> succ(n), root(n);
  std::vector<kascade::rank_t> w(n), rank(n);
  orc_gen_list(n, 123, 0, succ.data(), w.data());
  std::vector<kascade::idx_t> eroot = succ;
  std::vector<kascade::rank_t> erank(n);
  for (uint64_t i = 0; i < n; i++) erank[i] = succ[i] == i ? 0 : 1;
  orc_local_pointer_chasing(eroot.data(), erank.data(), n);
  // every rounds / base-algorithm combination gives the same (unique) result
  for (std::size_t rounds : {0, 1, 2, 5}) {
    for (auto base : {kascade::Algorithm::PointerDoubling, kascade::Algorithm::GatherChase,
                      kascade::Algorithm::AsyncPointerDoubling, kascade::Algorithm::RMAPointerDoubling}) {
      auto dist = kascade::set_initial_ranking_state(succ, root, rank, comm);
      EXPECT(dist.get_local_size() == n);
      kascade::SparseRulingSetConfig c = srs;
      c.sparse_ruling_set_rounds = rounds;
      c.base_algorithm = base;
      kascade::sparse_ruling_set(c, root, rank, dist, comm, grid);
      EXPECT(root == eroot && rank == erank);
    }
  }
  auto dist = kascade::set_initial_ranking_state(succ, root, rank, comm);
  kascade::pointer_doubling(pd, root, rank, dist, comm, grid);
  EXPECT(root == eroot && rank == erank);
  // a locality-biased list through the contraction switches: forced on for
  // the ruling set, off for doubling; the results do not change
  {
    std::vector<kascade::idx_t> ls(n), lroot(n);
    std::vector<kascade::rank_t> lw(n), lrank(n);
    orc_gen_list_local(n, 77, 0, ls.data(), lw.data());
    std::vector<kascade::idx_t> er2 = ls;
    std::vector<kascade::rank_t> ek2(n);
    for (uint64_t i = 0; i < n; i++) ek2[i] = ls[i] == i ? 0 : 1;
    orc_local_pointer_chasing(er2.data(), ek2.data(), n);
    for (bool contract : {true, false}) {
      auto d2 = kascade::set_initial_ranking_state(ls, lroot, lrank, comm);
      kascade::SparseRulingSetConfig c = srs;
      c.use_local_contraction = contract;
      kascade::sparse_ruling_set(c, lroot, lrank, d2, comm, grid);
      EXPECT(lroot == er2 && lrank == ek2);
      d2 = kascade::set_initial_ranking_state(ls, lroot, lrank, comm);
      kascade::PointerDoublingConfig p2 = pd;
      p2.use_local_preprocessing = contract;
      kascade::pointer_doubling(p2, lroot, lrank, d2, comm, grid);
      EXPECT(lroot == er2 && lrank == ek2);
    }
  }
  // the legacy base algorithm is promoted by the adapter
  kascade::SparseRulingSetConfig legacy;
  legacy.base_algorithm = kascade::Algorithm::SparseRulingSet;
  kascade::SparseRulingSet adapter(legacy);
  adapter.ingest(succ);
  adapter.run();
  EXPECT(adapter.get_root_array() == eroot && adapter.get_rank_array() == erank);
  // a base algorithm the reference rejects throws
  kascade::SparseRulingSetConfig badc;
  badc.base_algorithm = kascade::Algorithm::MPLR;
  bool thrown = false;
  try {
    kascade::sparse_ruling_set(badc, root, rank, dist, comm, grid);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  EXPECT(thrown);
  // more than one process: the whole array is not here
  FakeComm two{2};
  thrown = false;
  try {
    kascade::sparse_ruling_set(srs, root, rank, dist, two, grid);
  } catch (const kascade_amd::Error& e) {
    thrown = e.status == KC_EINVAL;
  }
  EXPECT(thrown);
  // conversions from a reference-typed configuration keep every member
  RefSrsConfig ref;
  ref.sparse_ruling_set_rounds = 3;
  ref.spawn = true;
  ref.ruler_selection = kascade::RulerSelectionStrategy::heuristic;
  auto conv = kascade_amd::from_reference(static_cast<kascade_amd::SparseRulingSetConfig const&>(ref));
  EXPECT(conv.sparse_ruling_set_rounds == 3 && conv.spawn &&
         conv.ruler_selection == kascade::RulerSelectionStrategy::heuristic);
  auto pconv = kascade_amd::from_reference(pd);
  EXPECT(pconv.aggregation_level == kascade::AggregationLevel::local && pconv.fallback_allgather_size_ratio == 3.0);
  auto econv = kascade_amd::from_reference<kascade_amd::SparseRulingSetConfig>(et);
  EXPECT(econv.algorithm == kascade::Algorithm::SparseRulingSet && econv.algo_config.has_value());

  // tree rooting through the Euler tour with the SRS config in algo_config,
  // and compute_euler_tour in the reference's order and length
  const uint64_t nt = 20011;
  std::vector<kascade::idx_t> parent(nt), troot(nt);
  std::vector<kascade::rank_t> tdepth(nt);
  orc_gen_tree(nt, 77, parent.data());
  std::vector<kascade::idx_t> er2 = parent;
  std::vector<kascade::rank_t> ed2(nt);
  orc_euler_tour_rooting(er2.data(), ed2.data(), nt);
  auto tdist = kascade::set_initial_ranking_state(parent, troot, tdepth, comm);
  kascade::rank_via_euler_tour(et, troot, tdepth, tdist, comm, grid);
  EXPECT(troot == er2 && tdepth == ed2);
  auto tour = kascade::compute_euler_tour(parent);
  std::vector<uint64_t> etour(2 * nt);
  const int64_t len = orc_compute_euler_tour(parent.data(), nt, etour.data(), nullptr);
  etour.resize(static_cast<size_t>(len));
  EXPECT(tour == etour);
}

int main() {
  reference_shaped_calls();
  pack_works();
  utilities_match_oracle();
  pack_sanity_check();
  adapters_match_gather_rank();
  errors_throw();
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("dropin_test: all checks passed\n");
  return 0;
}
