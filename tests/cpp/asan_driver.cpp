// Host-side AddressSanitizer exercise of the C ABI: the library's host code
// (contexts, scratch arena and chunk cache, staging, graph capture and replay,
// the serving loop with four submitting threads) instrumented, device code
// untouched.  Build with tests/cpp/build_asan.sh; run on a GPU with
//   ASAN_OPTIONS=detect_leaks=0:protect_shadow_gap=0 tests/cpp/_asan/asan_driver
// It prints each step and "asan driver done"; any ASan report is a failure.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <unistd.h>
#include <vector>
#include "kascade_amd.h"
#define CK(x) do { kc_status s_ = (x); if (s_ != KC_OK) { printf("FAIL %s -> %d at %d\n", #x, (int)s_, __LINE__); exit(1);} } while (0)
#define HK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e_), __LINE__); exit(1);} } while (0)
int main() {
  void* ctx = nullptr;
  CK(kc_context_create(0, 0, &ctx));
  printf("step lists\n"); fflush(stdout);
  for (uint64_t n : {1ull, 100ull, 100003ull, 1ull << 20, (1ull << 22) + 3}) {
    uint64_t *s, *r; int64_t *w, *k;
    HK(hipMalloc(&s, n * 8)); HK(hipMalloc(&w, n * 8)); HK(hipMalloc(&r, n * 8)); HK(hipMalloc(&k, n * 8));
    CK(kc_gen_list(ctx, n, 7 + n, 1, 0, s, w, nullptr));
    for (int algo : {KC_ALGO_AUTO, KC_ALGO_POINTER_DOUBLING, KC_ALGO_SPARSE_RULING_SET}) {
      for (int launch : {0, 1, 2}) {
        kc_rank_cfg cfg{}; cfg.algorithm = algo; cfg.launch = launch;
        for (int rep = 0; rep < 3; rep++) CK(kc_rank_list(ctx, s, w, n, r, k, &cfg, nullptr));
      }
    }
    // host arrays
    std::vector<uint64_t> hs(n), hr(n); std::vector<int64_t> hw(n), hk(n);
    HK(hipMemcpy(hs.data(), s, n * 8, hipMemcpyDeviceToHost)); HK(hipMemcpy(hw.data(), w, n * 8, hipMemcpyDeviceToHost));
    CK(kc_rank_list(ctx, hs.data(), hw.data(), n, hr.data(), hk.data(), nullptr, nullptr));
    HK(hipFree(s)); HK(hipFree(w)); HK(hipFree(r)); HK(hipFree(k));
  }
  printf("step errors\n"); fflush(stdout);
  {
    std::vector<uint64_t> cyc = {1, 2, 0, 3}; std::vector<uint64_t> ro(4); std::vector<int64_t> ra(4);
    kc_status st = kc_rank_list(ctx, cyc.data(), nullptr, 4, ro.data(), ra.data(), nullptr, nullptr);
    printf("cycle status %d\n", (int)st);
    std::vector<uint64_t> bad = {1, 9, 2};
    st = kc_rank_list(ctx, bad.data(), nullptr, 3, ro.data(), ra.data(), nullptr, nullptr);
    printf("range status %d\n", (int)st);
    std::vector<uint64_t> big(300000);
    for (uint64_t i = 0; i < big.size(); i++) big[i] = (i + 1) % big.size();  // one big cycle
    std::vector<uint64_t> bro(big.size()); std::vector<int64_t> bra(big.size());
    st = kc_rank_list(ctx, big.data(), nullptr, big.size(), bro.data(), bra.data(), nullptr, nullptr);
    printf("big cycle status %d\n", (int)st);
  }
  printf("step trees\n"); fflush(stdout);
  for (uint64_t n : {1ull, 3ull, 5000ull, 1ull << 20, (1ull << 21) + 5}) {
    uint64_t *p, *r; int64_t *d;
    HK(hipMalloc(&p, n * 8)); HK(hipMalloc(&r, n * 8)); HK(hipMalloc(&d, n * 8));
    CK(kc_gen_tree(ctx, n, 11 + n, p, nullptr));
    for (int tb : {0, 1, 2}) { kc_tree_cfg tc{}; tc.tour_build = tb; CK(kc_root_tree(ctx, p, n, r, d, &tc, nullptr)); }
    HK(hipFree(p)); HK(hipFree(r)); HK(hipFree(d));
  }
  {
    // a star: the linked build's grouped child insertion
    const uint64_t n = 300000;
    std::vector<uint64_t> hp(n, 7), hr(n); std::vector<int64_t> hd(n);
    CK(kc_root_tree(ctx, hp.data(), n, hr.data(), hd.data(), nullptr, nullptr));
  }
  printf("step segmented\n"); fflush(stdout);
  {
    const uint64_t nseg = 500;
    std::vector<uint64_t> off(nseg + 1, 0), seeds(nseg); std::vector<uint8_t> loc(nseg);
    for (uint64_t sidx = 0; sidx < nseg; sidx++) { off[sidx + 1] = off[sidx] + (sidx * 37 % 3000); seeds[sidx] = 5 + sidx; loc[sidx] = sidx % 3 == 0; }
    const uint64_t total = off[nseg];
    uint64_t *doff, *dseeds; uint8_t* dloc; uint32_t *ds, *dr; int64_t *dw, *dk;
    HK(hipMalloc(&doff, (nseg + 1) * 8)); HK(hipMalloc(&dseeds, nseg * 8)); HK(hipMalloc(&dloc, nseg));
    HK(hipMalloc(&ds, total * 4)); HK(hipMalloc(&dr, total * 4)); HK(hipMalloc(&dw, total * 8)); HK(hipMalloc(&dk, total * 8));
    HK(hipMemcpy(doff, off.data(), (nseg + 1) * 8, hipMemcpyHostToDevice));
    HK(hipMemcpy(dseeds, seeds.data(), nseg * 8, hipMemcpyHostToDevice));
    HK(hipMemcpy(dloc, loc.data(), nseg, hipMemcpyHostToDevice));
    CK(kc_gen_lists_segmented(ctx, doff, nseg, dseeds, dloc, 1, ds, dw, nullptr));
    for (int launch : {0, 1, 2}) { kc_rank_cfg cfg{}; cfg.launch = launch; CK(kc_rank_lists_segmented(ctx, ds, doff, nseg, dw, dr, dk, &cfg, nullptr)); }
    HK(hipFree(doff)); HK(hipFree(dseeds)); HK(hipFree(dloc)); HK(hipFree(ds)); HK(hipFree(dr)); HK(hipFree(dw)); HK(hipFree(dk));
  }
  printf("step segmented narrow\n"); fflush(stdout);
  {
    const uint64_t nseg = 1500;
    std::vector<uint64_t> off(nseg + 1, 0), seeds(nseg);
    for (uint64_t sidx = 0; sidx < nseg; sidx++) { off[sidx + 1] = off[sidx] + (sidx * 37 % 6000); seeds[sidx] = 9 + sidx; }
    const uint64_t total = off[nseg];
    uint64_t *doff, *dseeds; uint32_t *ds, *dr; int64_t *dw, *dk;
    HK(hipMalloc(&doff, (nseg + 1) * 8)); HK(hipMalloc(&dseeds, nseg * 8));
    HK(hipMalloc(&ds, total * 4)); HK(hipMalloc(&dr, total * 4)); HK(hipMalloc(&dw, total * 8)); HK(hipMalloc(&dk, total * 8));
    HK(hipMemcpy(doff, off.data(), (nseg + 1) * 8, hipMemcpyHostToDevice));
    HK(hipMemcpy(dseeds, seeds.data(), nseg * 8, hipMemcpyHostToDevice));
    CK(kc_gen_lists_segmented(ctx, doff, nseg, dseeds, nullptr, 1, ds, dw, nullptr));
    CK(kc_rank_lists_segmented(ctx, ds, doff, nseg, dw, dr, dk, nullptr, nullptr));
    HK(hipFree(doff)); HK(hipFree(dseeds)); HK(hipFree(ds)); HK(hipFree(dr)); HK(hipFree(dw)); HK(hipFree(dk));
  }
  printf("step server\n"); fflush(stdout);
  for (int mode = 0; mode < 2; mode++) {
    kc_server_cfg sc{}; sc.mem_budget_bytes = 4ull << 30; sc.one_job_per_launch = mode;
    void* sv = nullptr;
    CK(kc_server_create(ctx, &sc, &sv));
    const int jobs = 120;
    std::vector<uint64_t*> S(jobs), R(jobs); std::vector<int64_t*> W(jobs), K(jobs); std::vector<uint64_t> N(jobs);
    for (int j = 0; j < jobs; j++) {
      N[j] = 64 + (uint64_t)(j * 7919) % 200000;
      HK(hipMalloc(&S[j], N[j] * 8)); HK(hipMalloc(&R[j], N[j] * 8)); HK(hipMalloc(&W[j], N[j] * 8)); HK(hipMalloc(&K[j], N[j] * 8));
      if (j % 3) CK(kc_gen_list(ctx, N[j], 100 + j, 1, 0, S[j], W[j], nullptr));
      else CK(kc_gen_tree(ctx, N[j], 100 + j, S[j], nullptr));
    }
    HK(hipDeviceSynchronize());
    std::vector<std::thread> th;
    for (int t = 0; t < 4; t++) th.emplace_back([&, t] {
      std::vector<uint64_t> tickets;
      for (int j = t; j < jobs; j += 4) {
        kc_job job{}; job.kind = (j % 3) ? KC_JOB_LIST : KC_JOB_TREE; job.n = N[j]; job.succ = S[j];
        job.weight = (j % 3) ? W[j] : nullptr; job.root_out = R[j]; job.rank_out = K[j];
        uint64_t tk = 0; CK(kc_submit(sv, &job, &tk)); tickets.push_back(tk);
      }
      for (uint64_t tk : tickets) CK(kc_wait(sv, tk));
    });
    for (auto& x : th) x.join();
    // host-array jobs
    std::vector<std::vector<uint64_t>> hs(20), hr(20); std::vector<std::vector<int64_t>> hk(20);
    std::vector<uint64_t> tk(20);
    for (int j = 0; j < 20; j++) {
      const uint64_t n = 1000 + j * 500; hs[j].resize(n); hr[j].resize(n); hk[j].resize(n);
      for (uint64_t i = 0; i < n; i++) hs[j][i] = i + 1 < n ? i + 1 : i;
      kc_job job{}; job.kind = KC_JOB_LIST; job.n = n; job.succ = hs[j].data(); job.root_out = hr[j].data(); job.rank_out = hk[j].data();
      CK(kc_submit(sv, &job, &tk[j]));
    }
    for (int j = 0; j < 20; j++) CK(kc_wait(sv, tk[j]));
    kc_server_stats st{}; CK(kc_server_stats_get(sv, &st));
    kc_server_destroy(sv);
    for (int j = 0; j < jobs; j++) { HK(hipFree(S[j])); HK(hipFree(R[j])); HK(hipFree(W[j])); HK(hipFree(K[j])); }
  }
  printf("step contexts\n"); fflush(stdout);
  for (int c = 0; c < 4; c++) {
    void* c2 = nullptr; CK(kc_context_create(0, 0, &c2));
    std::vector<uint64_t> hs(50000), hr(50000); std::vector<int64_t> hk(50000);
    for (uint64_t i = 0; i < hs.size(); i++) hs[i] = (i + 7) < hs.size() ? i + 7 : i;
    CK(kc_rank_list(c2, hs.data(), nullptr, hs.size(), hr.data(), hk.data(), nullptr, nullptr));
    CK(kc_context_trim(c2));
    CK(kc_rank_list(c2, hs.data(), nullptr, hs.size(), hr.data(), hk.data(), nullptr, nullptr));
    kc_context_destroy(c2);
  }
  CK(kc_release_cached_scratch(0));
  kc_context_destroy(ctx);
  printf("asan driver done\n");
  fflush(stdout);
  _exit(0);  // the HIP runtime's own teardown under ASan is not what is checked here
}
