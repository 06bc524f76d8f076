// Self-checking C++ test of the drop-in header (include/kascade_amd.hpp),
// written like the reference's tests/pack_test.cpp and its harness adapters:
// the pack known-answer test on one device, the identity sanity check, and
// every AbstractAlgorithm adapter verified element-wise against the oracle's
// sequential walk (the harness's GatherRank reference, bench/detail/
// verification.cpp:9-88).  Exit status 0 = all checks passed.
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "kascade_amd.hpp"

extern "C" {  // CPU oracle (oracle/oracle.c), the checker
void orc_gen_list(uint64_t n, uint64_t seed, int weighted, uint64_t* succ, int64_t* w);
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

int main() {
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
