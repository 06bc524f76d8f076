// kascade_runner: the benchmark harness loop of the reference
// (bench/runner.cpp:16-89) on one GPU, through the C ABI.
//
//   generate input -> input check and reference run (verify level > 0) ->
//   `iterations` x { ingest, ranking, verify, statistics } -> one JSON report
//
// The report has the shape of bench/detail/reporting.cpp:53-63,
// {"config", "stats", "timer", "counters"}, with one timer object per
// iteration whose phases sit at the paths the reference's evaluation reads
// (eval/Config.jl: ["root", "ranking", "max"], ["root", "ranking",
// "chase_ruler", "max"], ["root", "ranking", "base_case", ...]); times are in
// seconds.  Verification (bench/detail/verification.cpp:9-88) compares roots,
// then ranks, element by element against a reference run and stops at the
// first mismatch; without --verify-continue-on-mismatch the process exits
// with status 1.  The reference run is an independent algorithm on the same
// device (pointer doubling, or the sparse ruling set / Euler tour when
// pointer doubling is the algorithm under test): results are unique, so any
// two correct algorithms agree bit for bit.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "kascade_amd.h"

#ifndef KC_GIT_TAG
#define KC_GIT_TAG "unknown"
#endif

namespace {

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "[error] %s\n", msg.c_str());
  std::exit(2);
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

// ---- a minimal ordered JSON value --------------------------------------------
struct Json {
  enum Kind { Null, Bool, Int, Uint, Real, Str, Arr, Obj } kind = Null;
  bool b = false;
  int64_t i = 0;
  uint64_t u = 0;
  double d = 0;
  std::string s;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  static Json boolean(bool v) { Json j; j.kind = Bool; j.b = v; return j; }
  static Json integer(int64_t v) { Json j; j.kind = Int; j.i = v; return j; }
  static Json uinteger(uint64_t v) { Json j; j.kind = Uint; j.u = v; return j; }
  static Json real(double v) { Json j; j.kind = Real; j.d = v; return j; }
  static Json str(const std::string& v) { Json j; j.kind = Str; j.s = v; return j; }
  static Json array() { Json j; j.kind = Arr; return j; }
  static Json object() { Json j; j.kind = Obj; return j; }

  Json& operator[](const std::string& key) {
    if (kind != Obj) { kind = Obj; obj.clear(); }
    for (auto& kv : obj)
      if (kv.first == key) return kv.second;
    obj.emplace_back(key, Json{});
    return obj.back().second;
  }
  void push(Json v) {
    if (kind != Arr) { kind = Arr; arr.clear(); }
    arr.push_back(std::move(v));
  }
  bool has(const std::string& key) const {
    for (auto& kv : obj)
      if (kv.first == key) return true;
    return false;
  }

  void dump(std::ostream& o, int indent, int depth) const {
    auto nl = [&](int dd) {
      o << '\n';
      for (int q = 0; q < dd * indent; q++) o << ' ';
    };
    switch (kind) {
      case Null: o << "null"; break;
      case Bool: o << (b ? "true" : "false"); break;
      case Int: o << i; break;
      case Uint: o << u; break;
      case Real: {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.17g", d);
        o << buf;
        break;
      }
      case Str: {
        o << '"';
        for (char c : s) {
          if (c == '"' || c == '\\') o << '\\' << c;
          else if (c == '\n') o << "\\n";
          else if (static_cast<unsigned char>(c) < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", c);
            o << buf;
          } else o << c;
        }
        o << '"';
        break;
      }
      case Arr:
        o << '[';
        for (size_t k = 0; k < arr.size(); k++) {
          if (k) o << ',';
          nl(depth + 1);
          arr[k].dump(o, indent, depth + 1);
        }
        if (!arr.empty()) nl(depth);
        o << ']';
        break;
      case Obj:
        o << '{';
        for (size_t k = 0; k < obj.size(); k++) {
          if (k) o << ',';
          nl(depth + 1);
          o << '"' << obj[k].first << "\": ";
          obj[k].second.dump(o, indent, depth + 1);
        }
        if (!obj.empty()) nl(depth);
        o << '}';
        break;
    }
  }
};

// ---- configuration (the reference's Config, bench/detail/benchmark_config.hpp) -
struct Config {
  std::string output_path = "stdout";
  uint64_t iterations = 1;
  std::string input = "list";  // list | list-local | forest | tree | gnm | rmat | rgg2d
  std::string input_processing = "bfs";  // graph inputs: bfs | eulertour
  uint64_t n = uint64_t{1} << 20;
  uint64_t m = uint64_t{1} << 22;  // graph inputs: edges
  uint64_t seed = 1;
  uint64_t lists = 1000;  // forest input
  bool weighted = false;
  std::string algorithm = "SparseRulingSet";  // PointerDoubling | SparseRulingSet | EulerTour
  uint32_t ruler_spacing = 0;
  uint32_t ruler_spacing_deep = 0;
  uint32_t rounds = 0;
  std::string tour_build = "auto";
  bool validate_list = false;
  uint64_t verify_level = 1;
  bool verify_continue_on_mismatch = false;
  std::string statistics_level = "none";  // none | basic | extensive | extensive-reduced-output
};

void usage() {
  std::printf(
      "usage: kascade_runner [options]\n"
      "  --algorithm PointerDoubling|SparseRulingSet|EulerTour   (default SparseRulingSet)\n"
      "  --input list|list-local|forest|tree|gnm|rmat|rgg2d      (default list)\n"
      "  --input-processing bfs|eulertour  (graph inputs: BFS forest, or its Euler tour)\n"
      "  --n N  --m M  --seed S  --lists L  --weighted\n"
      "  --iterations K                                          (default 1)\n"
      "  --verify-level L  --verify-continue-on-mismatch        (default 1)\n"
      "  --statistics none|basic|extensive|extensive-reduced-output\n"
      "  --ruler-spacing K  --ruler-spacing-deep K  --sparse-ruling-set-rounds R\n"
      "  --tour-build auto|sorted|linked  --validate-list\n"
      "  --output stdout|stderr|PATH[.json]\n");
}

Config parse(int argc, char** argv) {
  Config c;
  for (int k = 1; k < argc; k++) {
    std::string a = argv[k];
    auto val = [&]() -> std::string {
      if (k + 1 >= argc) die("missing value for " + a);
      return argv[++k];
    };
    auto num = [&]() -> uint64_t {
      std::string v = val();
      char* end = nullptr;
      uint64_t x = std::strtoull(v.c_str(), &end, 0);
      if (!end || *end) die("bad number for " + a + ": " + v);
      return x;
    };
    if (a == "-h" || a == "--help") { usage(); std::exit(0); }
    else if (a == "--output" || a == "-o") c.output_path = val();
    else if (a == "--iterations" || a == "-i") c.iterations = num();
    else if (a == "--input") c.input = val();
    else if (a == "--n") c.n = num();
    else if (a == "--m") c.m = num();
    else if (a == "--input-processing") c.input_processing = val();
    else if (a == "--seed") c.seed = num();
    else if (a == "--lists") c.lists = num();
    else if (a == "--weighted") c.weighted = true;
    else if (a == "--algorithm") c.algorithm = val();
    else if (a == "--ruler-spacing") c.ruler_spacing = static_cast<uint32_t>(num());
    else if (a == "--ruler-spacing-deep") c.ruler_spacing_deep = static_cast<uint32_t>(num());
    else if (a == "--sparse-ruling-set-rounds") c.rounds = static_cast<uint32_t>(num());
    else if (a == "--tour-build") c.tour_build = val();
    else if (a == "--validate-list") c.validate_list = true;
    else if (a == "--verify-level") c.verify_level = num();
    else if (a == "--verify-continue-on-mismatch") c.verify_continue_on_mismatch = true;
    else if (a == "--statistics") c.statistics_level = val();
    else die("unknown option " + a);
  }
  if (c.algorithm != "PointerDoubling" && c.algorithm != "SparseRulingSet" && c.algorithm != "EulerTour")
    die("invalid algorithm " + c.algorithm);
  if (c.input != "list" && c.input != "list-local" && c.input != "forest" && c.input != "tree" &&
      c.input != "gnm" && c.input != "rmat" && c.input != "rgg2d")
    die("invalid input " + c.input);
  if (c.input_processing != "bfs" && c.input_processing != "eulertour")
    die("invalid input processing " + c.input_processing);
  if (c.statistics_level != "none" && c.statistics_level != "basic" && c.statistics_level != "extensive" &&
      c.statistics_level != "extensive-reduced-output")
    die("invalid statistics level " + c.statistics_level);
  if (c.tour_build != "auto" && c.tour_build != "sorted" && c.tour_build != "linked")
    die("invalid tour build " + c.tour_build);
  if (c.n == 0) die("--n must be positive");
  if (c.input == "forest" && (c.lists == 0 || c.lists > c.n)) die("--lists must be in [1, n]");
  return c;
}

int stats_level_code(const std::string& s) {
  if (s == "basic") return KC_STATS_BASIC;
  if (s == "extensive-reduced-output") return KC_STATS_REDUCED_EXTENSIVE;
  if (s == "extensive") return KC_STATS_EXTENSIVE;
  return KC_STATS_NONE;
}

struct Device {
  void* ctx = nullptr;
  void check(kc_status s, const char* what) const {
    if (s != KC_OK) die(std::string(what) + " failed: " + kc_last_error(ctx));
  }
};

template <class T>
T* dalloc(uint64_t n) {
  void* p = nullptr;
  hip_check(hipMalloc(&p, std::max<uint64_t>(n, 1) * sizeof(T)), "hipMalloc");
  return static_cast<T*>(p);
}

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// The input as a successor (or parent) array on the host, generated on the
// device by the library's seeded generators (KaGen's generators are not
// available here; bench/input/generation.cpp:60-96 selects the input).
void generate(const Config& c, const Device& d, std::vector<uint64_t>& succ, std::vector<int64_t>& w) {
  const uint64_t n = c.n;
  uint64_t* ds = dalloc<uint64_t>(n);
  int64_t* dw = dalloc<int64_t>(n);
  succ.assign(n, 0);
  w.assign(n, 0);
  if (c.input == "tree") {
    d.check(kc_gen_tree(d.ctx, n, c.seed, ds, nullptr), "kc_gen_tree");
    hip_check(hipMemcpy(succ.data(), ds, n * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy");
  } else if (c.input == "gnm" || c.input == "rmat" || c.input == "rgg2d") {
    // generate_input with "bfs" / "eulertour" processing (generation.cpp:74-90)
    uint64_t m = c.m;
    if (c.input == "rgg2d") d.check(kc_gen_rgg2d(d.ctx, n, c.m, c.seed, nullptr, nullptr, 0, &m, nullptr), "kc_gen_rgg2d");
    uint64_t* src = dalloc<uint64_t>(m ? m : 1);
    uint64_t* dst = dalloc<uint64_t>(m ? m : 1);
    uint64_t* xadj = dalloc<uint64_t>(n + 1);
    uint64_t* adj = dalloc<uint64_t>(2 * m ? 2 * m : 1);
    if (c.input == "rgg2d")
      d.check(kc_gen_rgg2d(d.ctx, n, c.m, c.seed, src, dst, m, &m, nullptr), "kc_gen_rgg2d");
    else
      d.check(kc_gen_graph(d.ctx, c.input == "gnm" ? KC_GRAPH_GNM : KC_GRAPH_RMAT, n, m, c.seed, src, dst, nullptr),
              "kc_gen_graph");
    d.check(kc_build_csr(d.ctx, src, dst, n, m, xadj, adj, nullptr), "kc_build_csr");
    d.check(kc_bfs_forest(d.ctx, xadj, adj, n, ds, nullptr), "kc_bfs_forest");
    (void)hipFree(src);
    (void)hipFree(dst);
    (void)hipFree(xadj);
    (void)hipFree(adj);
    if (c.input_processing == "eulertour") {
      // compute_euler_tour (eulertour.cpp:379-392): the tour is the input list
      uint64_t* tour = dalloc<uint64_t>(2 * n);
      uint64_t len = 0;
      d.check(kc_compute_euler_tour(d.ctx, ds, n, tour, nullptr, &len, nullptr), "kc_compute_euler_tour");
      succ.assign(len, 0);
      w.assign(len, 0);
      hip_check(hipMemcpy(succ.data(), tour, len * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy");
      (void)hipFree(tour);
    } else {
      hip_check(hipMemcpy(succ.data(), ds, n * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy");
    }
  } else if (c.input == "forest") {
    // `lists` lists of near-equal sizes laid out one after another
    std::vector<uint64_t> off(c.lists + 1), seeds(c.lists);
    for (uint64_t s = 0; s <= c.lists; s++) off[s] = n * s / c.lists;
    for (uint64_t s = 0; s < c.lists; s++) seeds[s] = c.seed * 1000003 + s;
    std::vector<uint8_t> local(c.lists, 0);
    uint64_t* doff = dalloc<uint64_t>(c.lists + 1);
    uint64_t* dseeds = dalloc<uint64_t>(c.lists);
    uint8_t* dlocal = dalloc<uint8_t>(c.lists);
    uint32_t* dsucc32 = dalloc<uint32_t>(n);
    hip_check(hipMemcpy(doff, off.data(), off.size() * 8, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(dseeds, seeds.data(), seeds.size() * 8, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(dlocal, local.data(), local.size(), hipMemcpyHostToDevice), "copy");
    d.check(kc_gen_lists_segmented(d.ctx, doff, c.lists, dseeds, dlocal, c.weighted ? 1 : 0, dsucc32,
                                   c.weighted ? dw : nullptr, nullptr),
            "kc_gen_lists_segmented");
    std::vector<uint32_t> s32(n);
    hip_check(hipMemcpy(s32.data(), dsucc32, n * 4, hipMemcpyDeviceToHost), "copy");
    for (uint64_t s = 0; s < c.lists; s++)
      for (uint64_t i = off[s]; i < off[s + 1]; i++) succ[i] = off[s] + s32[i];
    if (c.weighted) hip_check(hipMemcpy(w.data(), dw, n * 8, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(doff);
    (void)hipFree(dseeds);
    (void)hipFree(dlocal);
    (void)hipFree(dsucc32);
  } else {
    d.check(kc_gen_list(d.ctx, n, c.seed, c.weighted ? 1 : 0, c.input == "list-local" ? 1 : 0, ds,
                        c.weighted ? dw : nullptr, nullptr),
            "kc_gen_list");
    hip_check(hipMemcpy(succ.data(), ds, n * sizeof(uint64_t), hipMemcpyDeviceToHost), "copy");
    if (c.weighted) hip_check(hipMemcpy(w.data(), dw, n * 8, hipMemcpyDeviceToHost), "copy");
  }
  if (!c.weighted || c.input == "tree" || c.input == "gnm" || c.input == "rmat" || c.input == "rgg2d")
    // unit weights, as set_initial_ranking_state
    for (uint64_t i = 0; i < succ.size(); i++) w[i] = succ[i] == i ? 0 : 1;
  (void)hipFree(ds);
  (void)hipFree(dw);
}

// One ranking, in place on device arrays (root in = successors, rank in =
// distances), with the library's phase profile enabled.
void rank_in_place(const Device& d, const std::string& algo, const Config& c, uint64_t* root, int64_t* rank,
                   uint64_t n) {
  kc_rank_cfg rc{};
  rc.ruler_spacing = c.ruler_spacing;
  rc.ruler_spacing_deep = c.ruler_spacing_deep;
  rc.max_levels = c.rounds;
  // the input check of the algorithm under test (not of the reference run)
  rc.validate = c.validate_list && algo == c.algorithm && algo != "EulerTour" ? 1u : 0u;
  if (algo == "EulerTour") {
    kc_tree_cfg tc{};
    tc.rank = rc;
    tc.rank.algorithm = KC_ALGO_SPARSE_RULING_SET;
    tc.tour_build = c.tour_build == "sorted" ? KC_TOUR_SORTED : c.tour_build == "linked" ? KC_TOUR_LINKED
                                                                                          : KC_TOUR_AUTO;
    d.check(kc_rank_via_euler_tour(d.ctx, &tc, root, rank, n, nullptr), "kc_rank_via_euler_tour");
  } else if (algo == "PointerDoubling") {
    rc.algorithm = KC_ALGO_POINTER_DOUBLING;
    d.check(kc_pointer_doubling(d.ctx, &rc, root, rank, n, nullptr), "kc_pointer_doubling");
  } else {
    rc.algorithm = KC_ALGO_SPARSE_RULING_SET;
    d.check(kc_sparse_ruling_set(d.ctx, &rc, root, rank, n, nullptr), "kc_sparse_ruling_set");
  }
}

// Phase name of the library profile -> path under root.ranking (level L of
// the recursion is L nested "base_case" objects, as the reference's timer
// nests the recursive calls).
std::vector<std::string> phase_path(const std::string& name) {
  std::vector<std::string> path = {"root", "ranking"};
  auto level_of = [&](const std::string& prefix) -> int {
    if (name.rfind(prefix, 0) != 0) return -1;
    return std::atoi(name.c_str() + prefix.size());
  };
  int l;
  if ((l = level_of("chase_ruler_L")) >= 0) {
    for (int q = 0; q < l; q++) path.push_back("base_case");
    path.push_back("chase_ruler");
  } else if ((l = level_of("ruler_propagation_L")) >= 0) {
    for (int q = 0; q < l; q++) path.push_back("base_case");
    path.push_back("ruler_propagation");
  } else if (name == "unclaimed") {
    path.push_back("fixup_unreached");
  } else if (name == "pointer_doubling" || name == "small_rank") {
    path.push_back("base_case");
    path.push_back(name);
  } else {
    path.push_back(name);
  }
  return path;
}

void add_time(Json& timer, const std::vector<std::string>& path, double seconds) {
  Json* j = &timer;
  for (const auto& key : path) j = &(*j)[key];
  Json& mx = (*j)["max"];
  if (mx.kind != Json::Arr) mx = Json::array();
  mx.push(Json::real(seconds));
}

Json stats_json(const Device& d, const Config& c, const uint64_t* root, const int64_t* rank, uint64_t n) {
  const int level = stats_level_code(c.statistics_level);
  Json out = Json::object();
  if (level == KC_STATS_NONE) return out;
  kc_stats st{};
  std::vector<uint64_t> tr, ts;
  std::vector<int64_t> tm, tsum;
  if (level == KC_STATS_EXTENSIVE) {
    tr.resize(n);
    ts.resize(n);
    tm.resize(n);
    tsum.resize(n);
  }
  d.check(kc_compute_stats(d.ctx, level, root, rank, n, &st, tr.data(), ts.data(), tm.data(), tsum.data(),
                           tr.size(), nullptr),
          "kc_compute_stats");
  Json& b = out["basic_stats"];
  b["num_trees"] = Json::uinteger(st.num_trees);
  b["max_rank"] = Json::integer(st.max_rank);
  b["avg_rank"] = Json::real(st.avg_rank);
  if (level == KC_STATS_BASIC) return out;
  // the serialized fields of ExtensiveStats (bench/detail/serialization.hpp:142-150)
  Json& e = out["extensive_stats"];
  e["num_nontrivial_trees"] = Json::uinteger(st.num_nontrivial_trees);
  e["max_size"] = Json::uinteger(st.max_size);
  e["avg_size"] = Json::real(st.avg_size);
  e["nontrivial_avg_size"] = Json::real(st.nontrivial_avg_size);
  e["nontrivial_avg_rank"] = Json::real(st.nontrivial_avg_rank);
  Json& per = e["per_tree_stats"];
  if (level == KC_STATS_EXTENSIVE) {
    per = Json::array();
    for (uint64_t k = 0; k < st.num_nontrivial_trees && k < tr.size(); k++) {
      Json t = Json::object();
      t["size"] = Json::uinteger(ts[k]);
      t["max_rank"] = Json::integer(tm[k]);
      t["rank_sum"] = Json::integer(tsum[k]);
      per.push(std::move(t));
    }
  }
  return out;
}

// verification.cpp:9-88: roots first, then ranks; stop at the first mismatch
bool verify(const Config& c, const std::vector<uint64_t>& root, const std::vector<int64_t>& rank,
            const std::vector<uint64_t>& ref_root, const std::vector<int64_t>& ref_rank) {
  if (c.verify_level == 0) return true;
  auto fail = [&](const std::string& msg) {
    std::fprintf(stderr, "[error] %s\n", msg.c_str());
    if (!c.verify_continue_on_mismatch) std::exit(1);
  };
  bool ok = true;
  if (root.size() != ref_root.size()) fail("Root array size does not match reference");
  for (uint64_t i = 0; i < root.size() && i < ref_root.size(); i++) {
    if (root[i] != ref_root[i]) {
      ok = false;
      fail("Root array does not match reference: " + std::to_string(root[i]) + " != " +
           std::to_string(ref_root[i]) + " at index " + std::to_string(i));
      break;
    }
  }
  std::fprintf(stderr, ok ? "[info] Root array matches reference.\n" : "[error] Root array does not match reference.\n");
  bool ok_rank = true;
  if (rank.size() != ref_rank.size()) fail("Rank array size does not match reference");
  for (uint64_t i = 0; i < rank.size() && i < ref_rank.size(); i++) {
    if (rank[i] != ref_rank[i]) {
      ok_rank = false;
      fail("Rank array does not match reference: " + std::to_string(rank[i]) + " != " +
           std::to_string(ref_rank[i]) + " at index " + std::to_string(i));
      break;
    }
  }
  std::fprintf(stderr, ok_rank ? "[info] Rank array matches reference.\n" : "[error] Rank array does not match reference.\n");
  return ok && ok_rank;
}

std::unique_ptr<std::ostream> open_output(const std::string& path_in) {
  // make_output_stream (reporting.cpp:14-31)
  if (path_in == "stdout") return std::make_unique<std::ostream>(std::cout.rdbuf());
  if (path_in == "stderr") return std::make_unique<std::ostream>(std::cerr.rdbuf());
  std::filesystem::path path(path_in);
  if (path.extension() != ".json") path.replace_extension(".json");
  auto f = std::make_unique<std::ofstream>(path);
  if (!f->is_open()) die("Failed to open output file: " + path.string());
  return f;
}

}  // namespace

int main(int argc, char** argv) {
  const Config c = parse(argc, argv);
  Device d;
  if (kc_context_create(0, 0, &d.ctx) != KC_OK) die("kc_context_create failed");
  std::vector<uint64_t> succ;
  std::vector<int64_t> w;
  generate(c, d, succ, w);
  const uint64_t n = succ.size();  // 2 x the vertices for an Euler tour input

  uint64_t* d_root = dalloc<uint64_t>(n);
  int64_t* d_rank = dalloc<int64_t>(n);
  Json report = Json::object();
  Json stats = Json::object();
  // one rank holds the whole input (get_input_distribution_stats)
  for (const char* key : {"input_distribution", "rebalanced_input_distribution"}) {
    Json& s = stats[key];
    s["min_local_size"] = Json::uinteger(n);
    s["max_local_size"] = Json::uinteger(n);
    s["total_size"] = Json::uinteger(n);
  }

  // reference run (runner.cpp:41-55)
  std::vector<uint64_t> ref_root;
  std::vector<int64_t> ref_rank;
  Json pre_timer = Json::object();
  std::string ref_algo;
  if (c.verify_level > 0) {
    auto t0 = Clock::now();
    int32_t is_list = 0;
    hip_check(hipMemcpy(d_root, succ.data(), n * 8, hipMemcpyHostToDevice), "copy");
    d.check(kc_is_list(d.ctx, d_root, n, &is_list, nullptr), "kc_is_list");
    add_time(pre_timer, {"root", "list_check"}, since(t0));
    std::fprintf(stderr, "[info] Input is list? %s\n", is_list ? "true" : "false");
    if (c.algorithm == "PointerDoubling") ref_algo = is_list ? "SparseRulingSet" : "EulerTour";
    else ref_algo = "PointerDoubling";
    t0 = Clock::now();
    hip_check(hipMemcpy(d_rank, w.data(), n * 8, hipMemcpyHostToDevice), "copy");
    rank_in_place(d, ref_algo, c, d_root, d_rank, n);
    add_time(pre_timer, {"root", "reference_run"}, since(t0));
    ref_root.resize(n);
    ref_rank.resize(n);
    hip_check(hipMemcpy(ref_root.data(), d_root, n * 8, hipMemcpyDeviceToHost), "copy");
    hip_check(hipMemcpy(ref_rank.data(), d_rank, n * 8, hipMemcpyDeviceToHost), "copy");
    std::fprintf(stderr, "[info] Reference (%s): finished.\n", ref_algo.c_str());
  }

  Json timers = Json::array();
  Json counters = Json::array();
  std::vector<uint64_t> root(n);
  std::vector<int64_t> rank(n);
  bool all_ok = true;
  for (uint64_t it = 0; it < c.iterations; it++) {
    Json timer = it == 0 ? pre_timer : Json::object();
    // ingest (AbstractAlgorithm::ingest): the input into the algorithm's arrays
    hip_check(hipDeviceSynchronize(), "sync");
    auto t0 = Clock::now();
    hip_check(hipMemcpy(d_root, succ.data(), n * 8, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(d_rank, w.data(), n * 8, hipMemcpyHostToDevice), "copy");
    add_time(timer, {"root", "ingest_graph"}, since(t0));
    d.check(kc_profile_enable(d.ctx, 1), "kc_profile_enable");
    d.check(kc_profile_reset(d.ctx), "kc_profile_reset");
    d.check(kc_trace_enable(d.ctx, 1), "kc_trace_enable");
    t0 = Clock::now();
    rank_in_place(d, c.algorithm, c, d_root, d_rank, n);
    hip_check(hipDeviceSynchronize(), "sync");
    add_time(timer, {"root", "ranking"}, since(t0));
    const char* names[64];
    double ms[64];
    uint64_t counts[64];
    const uint32_t np = kc_profile_read(d.ctx, 64, names, ms, counts);
    for (uint32_t k = 0; k < np && k < 64; k++) add_time(timer, phase_path(names[k]), ms[k] * 1e-3);
    d.check(kc_profile_enable(d.ctx, 0), "kc_profile_enable");
    // RulerTrace counters (trace.hpp:31-119), one value per ruling-set level
    Json counter = Json::object();
    {
      std::vector<kc_trace_level> tr(kc_trace_read(d.ctx, nullptr, 0));
      kc_trace_read(d.ctx, tr.data(), static_cast<uint32_t>(tr.size()));
      d.check(kc_trace_enable(d.ctx, 0), "kc_trace_enable");
      auto add = [&](const char* name, std::initializer_list<const char*> modes, Json v) {
        for (const char* mode : modes) {
          Json& a = counter["root"][name][mode];
          if (a.kind != Json::Arr) a = Json::array();
          a.push(v);
        }
      };
      for (const auto& t : tr) {
        if (t.kind == KC_TRACE_PD) add("pointer_doubling_rounds", {"max"}, Json::uinteger(t.rounds));
        if (t.kind != KC_TRACE_SRS) continue;
        add("ruler_list_length_min", {"min"}, Json::uinteger(t.chain_length_min));
        add("ruler_list_length_max", {"max"}, Json::uinteger(t.chain_length_max));
        add("ruler_list_length_avg", {"max"}, Json::integer(static_cast<int64_t>(t.chain_length_avg)));
        add("local_num_ruler", {"max", "min"}, Json::uinteger(t.rulers));
        add("local_subproblem_size", {"sum", "max", "min"}, Json::uinteger(t.rulers + t.tail_rulers));
        add("ruler_chasing_rounds", {"max", "min"}, Json::uinteger(t.chain_length_max));
        add("num_unreached", {"sum", "max", "min"}, Json::uinteger(t.unreached));
        add("num_tail_rulers", {"sum"}, Json::uinteger(t.tail_rulers));
      }
    }
    hip_check(hipMemcpy(root.data(), d_root, n * 8, hipMemcpyDeviceToHost), "copy");
    hip_check(hipMemcpy(rank.data(), d_rank, n * 8, hipMemcpyDeviceToHost), "copy");
    std::fprintf(stderr, "[info] Finished run %llu in %.6f secs.\n", static_cast<unsigned long long>(it),
                 timer["root"]["ranking"]["max"].arr.back().d);
    all_ok &= verify(c, root, rank, ref_root, ref_rank);
    if (!stats.has("input_stats"))  // Report::push_stats keeps the first value
      stats["input_stats"] = stats_json(d, c, d_root, d_rank, n);
    timers.push(std::move(timer));
    counters.push(std::move(counter));
  }

  hipDeviceProp_t prop{};
  std::string device = "unknown";
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess) device = std::string(prop.name) + " " + prop.gcnArchName;
  Json& cfg = report["config"];
  cfg["num_ranks"] = Json::uinteger(1);
  cfg["device"] = Json::str(device);
  cfg["git_tag"] = Json::str(KC_GIT_TAG);
  cfg["output_path"] = Json::str(c.output_path);
  cfg["iterations"] = Json::uinteger(c.iterations);
  {
    std::ostringstream g;
    const bool graph = c.input == "gnm" || c.input == "rmat" || c.input == "rgg2d";
    g << c.input << ";n=" << c.n << ";seed=" << c.seed << ";weighted=" << (c.weighted ? 1 : 0);
    if (c.input == "forest") g << ";lists=" << c.lists;
    if (graph) g << ";m=" << c.m;
    cfg["input"]["kagen_option_string"] = Json::str(g.str());
    cfg["input"]["input_processing"] = Json::str(graph ? c.input_processing : "none");
  }
  cfg["algorithm"] = Json::str(c.algorithm);
  cfg["sparse_ruling_set"]["ruler_spacing"] = Json::uinteger(c.ruler_spacing);
  cfg["sparse_ruling_set"]["ruler_spacing_deep"] = Json::uinteger(c.ruler_spacing_deep);
  cfg["sparse_ruling_set"]["sparse_ruling_set_rounds"] = Json::uinteger(c.rounds);
  cfg["sparse_ruling_set"]["validate_list"] = Json::boolean(c.validate_list);
  cfg["euler_tour"]["algorithm"] = Json::str("SparseRulingSet");
  cfg["euler_tour"]["tour_build"] = Json::str(c.tour_build);
  cfg["verify_level"] = Json::uinteger(c.verify_level);
  cfg["verify_continue_on_mismatch"] = Json::boolean(c.verify_continue_on_mismatch);
  cfg["verification_reference"] = Json::str(ref_algo);
  cfg["statistics_level"] = Json::str(c.statistics_level);
  report["stats"] = std::move(stats);
  report["timer"] = std::move(timers);
  report["counters"] = std::move(counters);

  auto out = open_output(c.output_path);
  report.dump(*out, 2, 0);
  *out << '\n';
  out->flush();
  (void)hipFree(d_root);
  (void)hipFree(d_rank);
  kc_context_destroy(d.ctx);
  return all_ok ? 0 : 1;
}
