/*
 * oracle.c -- CPU restatement of KAscade's list-ranking and tree-rooting
 * algorithms.  TEST INFRASTRUCTURE ONLY: this file is the checker used by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  It is
 * never linked into, loaded by or called from the kascade_amd product library.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the niklas-uhl/kascade source tree).  What pins each function is listed in
 * oracle/README.md; tests/test_oracle.py holds the pinning checks.
 *
 * Conventions (lib/kascade/types.hpp:8-9, lib/kascade/bits.hpp:9-21):
 *   idx_t  = uint64_t  (successor / root ids, MSB reserved as root flag)
 *   rank_t = int64_t   (weighted distance to the list tail)
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ROOT_FLAG (UINT64_C(1) << 63) /* bits.hpp:9 root_flag_mask */

/* ------------------------------------------------------------------------ */
/* Seeded generators.  The same arithmetic is implemented on the device by  */
/* kascade_amd/csrc/gen.hip; tests check both bit-for-bit.                  */
/* ------------------------------------------------------------------------ */

uint64_t orc_splitmix64(uint64_t x) {
  x += UINT64_C(0x9E3779B97F4A7C15);
  x = (x ^ (x >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
  x = (x ^ (x >> 27)) * UINT64_C(0x94D049BB133111EB);
  return x ^ (x >> 31);
}

/* Balanced Feistel network over [0, 2^(2h)) with cycle walking into [0, n).
 * A seeded pseudo-random permutation that can be evaluated (and inverted)
 * per element, so the device can generate 2^27-element lists in one pass. */
#define FEISTEL_ROUNDS 4
static unsigned feistel_half_bits(uint64_t n) {
  unsigned b = 2;
  while (b < 64 && (UINT64_C(1) << b) < n) b++;
  return (b + 1) / 2;
}
static uint64_t feistel_fwd(uint64_t x, unsigned h, uint64_t seed) {
  uint64_t mask = (UINT64_C(1) << h) - 1;
  uint64_t l = x >> h, r = x & mask;
  for (int k = 0; k < FEISTEL_ROUNDS; k++) {
    uint64_t f = orc_splitmix64(r ^ orc_splitmix64(seed + (uint64_t)k)) & mask;
    uint64_t t = l ^ f;
    l = r;
    r = t;
  }
  return (l << h) | r;
}
static uint64_t feistel_inv(uint64_t y, unsigned h, uint64_t seed) {
  uint64_t mask = (UINT64_C(1) << h) - 1;
  uint64_t l = y >> h, r = y & mask;
  for (int k = FEISTEL_ROUNDS - 1; k >= 0; k--) {
    uint64_t f = orc_splitmix64(l ^ orc_splitmix64(seed + (uint64_t)k)) & mask;
    uint64_t t = r ^ f;
    r = l;
    l = t;
  }
  return (l << h) | r;
}
uint64_t orc_perm(uint64_t x, uint64_t n, uint64_t seed) {
  unsigned h = feistel_half_bits(n);
  do { x = feistel_fwd(x, h, seed); } while (x >= n);
  return x;
}
uint64_t orc_perm_inv(uint64_t y, uint64_t n, uint64_t seed) {
  unsigned h = feistel_half_bits(n);
  do { y = feistel_inv(y, h, seed); } while (y >= n);
  return y;
}

static int64_t gen_weight(uint64_t v, uint64_t wseed) {
  return 1 + (int64_t)(orc_splitmix64(wseed ^ (v * UINT64_C(0xD1B54A32D192ED03))) % 100);
}

/* Random list: order pi(i) = perm(i); succ[pi(i)] = pi(i+1); the tail points
 * to itself.  weighted: w uniform in [1,100] with w[tail] = 0 (the tail must
 * carry 0: pointer_doubling_generic.hpp:431); else unit weights with
 * w[tail] = 0, exactly set_initial_ranking_state (list_ranking.cpp:60-80). */
void orc_gen_list(uint64_t n, uint64_t seed, int weighted, uint64_t *succ, int64_t *w) {
  uint64_t wseed = orc_splitmix64(seed ^ UINT64_C(0x5EED));
  for (uint64_t v = 0; v < n; v++) {
    uint64_t i = orc_perm_inv(v, n, seed);
    int tail = (i + 1 == n);
    succ[v] = tail ? v : orc_perm(i + 1, n, seed);
    if (w) w[v] = tail ? 0 : (weighted ? gen_weight(v, wseed) : 1);
  }
}

/* Locality-biased list: ids cut into blocks of 10 consecutive ids; blocks are
 * visited in a seeded permuted order, ids inside a block ascending.  So
 * succ(v) = v+1 with probability 0.9, else a jump to a random block start. */
#define LOCAL_BLOCK 10
void orc_gen_list_local(uint64_t n, uint64_t seed, int weighted, uint64_t *succ, int64_t *w) {
  uint64_t nb = (n + LOCAL_BLOCK - 1) / LOCAL_BLOCK;
  uint64_t wseed = orc_splitmix64(seed ^ UINT64_C(0x5EED));
  for (uint64_t v = 0; v < n; v++) {
    uint64_t b = v / LOCAL_BLOCK;
    uint64_t bend = (b + 1) * LOCAL_BLOCK < n ? (b + 1) * LOCAL_BLOCK : n;
    int tail = 0;
    if (v + 1 < bend) {
      succ[v] = v + 1;
    } else {
      uint64_t q = orc_perm_inv(b, nb, seed);
      if (q + 1 == nb) {
        succ[v] = v;
        tail = 1;
      } else {
        succ[v] = orc_perm(q + 1, nb, seed) * LOCAL_BLOCK;
      }
    }
    if (w) w[v] = tail ? 0 : (weighted ? gen_weight(v, wseed) : 1);
  }
}

/* Closed form for the random list: rank[pi(i)] = sum_{j=i}^{n-2} w[pi(j)],
 * root = pi(n-1).  Independent of any ranking algorithm. */
void orc_list_closed_form(uint64_t n, uint64_t seed, const int64_t *w, uint64_t *root, int64_t *rank) {
  if (n == 0) return;
  uint64_t tail = orc_perm(n - 1, n, seed);
  int64_t acc = 0;
  for (uint64_t i = n; i-- > 0;) {
    uint64_t v = orc_perm(i, n, seed);
    acc += w[v];
    rank[v] = acc;
    root[v] = tail;
  }
}

/* Random recursive tree relabelled by sigma: generation parent p(i) uniform
 * in [0, i); parent'[sigma(i)] = sigma(p(i)); sigma(0) is the root. */
static uint64_t tree_gen_parent(uint64_t i, uint64_t pseed) {
  return orc_splitmix64(pseed ^ (i * UINT64_C(0x9E3779B97F4A7C15))) % i;
}
void orc_gen_tree(uint64_t n, uint64_t seed, uint64_t *parent) {
  uint64_t pseed = orc_splitmix64(seed ^ UINT64_C(0x7EE));
  for (uint64_t v = 0; v < n; v++) {
    uint64_t i = orc_perm_inv(v, n, seed);
    parent[v] = (i == 0) ? v : orc_perm(tree_gen_parent(i, pseed), n, seed);
  }
}
/* Closed form: depth_gen[i] = depth_gen[p(i)] + 1 in generation order. */
void orc_tree_closed_form(uint64_t n, uint64_t seed, uint64_t *root, int64_t *depth) {
  if (n == 0) return;
  uint64_t pseed = orc_splitmix64(seed ^ UINT64_C(0x7EE));
  int64_t *dg = (int64_t *)malloc(n * sizeof(int64_t));
  dg[0] = 0;
  for (uint64_t i = 1; i < n; i++) dg[i] = dg[tree_gen_parent(i, pseed)] + 1;
  uint64_t r = orc_perm(0, n, seed);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t v = orc_perm(i, n, seed);
    depth[v] = dg[i];
    root[v] = r;
  }
  free(dg);
}

/* ------------------------------------------------------------------------ */
/* Sequential walk -- lib/kascade/local_pointer_chasing.cpp:18-59 (the      */
/* identity-distribution overload :62-75).  In place: succ -> root,         */
/* rank (distance to successor) -> distance to root.  Handles forests.      */
/* ------------------------------------------------------------------------ */
void orc_local_pointer_chasing(uint64_t *succ, int64_t *rank, uint64_t n) {
  unsigned char *has_pred = (unsigned char *)calloc(n ? n : 1, 1);
  uint64_t *path = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; i++) has_pred[succ[i]] = 1; /* :23-31 */
  for (uint64_t leaf = 0; leaf < n; leaf++) {              /* :41 leaves */
    if (has_pred[leaf]) continue;
    uint64_t len = 0, cur = leaf;
    while (succ[cur] != cur) { /* :44-47 walk to root */
      path[len++] = cur;
      cur = succ[cur];
    }
    uint64_t root = cur;
    /* :49-54 assign back along the path, skipping the node next to root */
    for (uint64_t k = len; k-- > 0;) {
      if (k == len - 1) continue;
      uint64_t idx = path[k];
      rank[idx] += rank[succ[idx]];
      succ[idx] = root;
    }
  }
  free(path);
  free(has_pred);
}

/* list_ranking.cpp:60-80 -- root := succ, rank := 1, rank := 0 on roots. */
void orc_set_initial_ranking_state(const uint64_t *succ, uint64_t *root, int64_t *rank, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    root[i] = succ[i];
    rank[i] = (succ[i] == i) ? 0 : 1;
  }
}

/* ------------------------------------------------------------------------ */
/* Jacobi Wyllie rounds -- the state transition of one sync pointer-doubling */
/* step (pointer_doubling_generic.hpp:187-243 with update_via_read_offsets   */
/* :43-81): all replies are built from the state before the round.          */
/* succ entries carry the root flag once they point at a root (:418-438).    */
/* Runs `rounds` rounds (or to completion if rounds < 0); returns the number */
/* of rounds executed.  The caller strips the flags.                         */
/* ------------------------------------------------------------------------ */
int orc_pd_rounds(uint64_t *succ, int64_t *rank, uint64_t n, int rounds) {
  uint64_t *s2 = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
  int64_t *r2 = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
  for (uint64_t i = 0; i < n; i++)
    if (succ[i] == i) succ[i] |= ROOT_FLAG;
  int done_rounds = 0;
  for (;;) {
    if (rounds >= 0 && done_rounds >= rounds) break;
    uint64_t active = 0;
    for (uint64_t i = 0; i < n; i++) {
      uint64_t s = succ[i];
      if (s & ROOT_FLAG) {
        s2[i] = s;
        r2[i] = rank[i];
        continue;
      }
      s2[i] = succ[s];
      r2[i] = rank[i] + rank[s];
      if (!(s2[i] & ROOT_FLAG)) active++;
    }
    memcpy(succ, s2, n * sizeof(uint64_t));
    memcpy(rank, r2, n * sizeof(int64_t));
    done_rounds++;
    if (rounds < 0 && active == 0) break;
    if (done_rounds > 70) break; /* no list has more than 2^63 elements */
  }
  free(s2);
  free(r2);
  return done_rounds;
}

/* ------------------------------------------------------------------------ */
/* Sync pointer doubling on p emulated ranks (one thread per rank, a shared  */
/* memory all-to-all).  Restates pointer_doubling_generic.hpp:608-700 with   */
/* the submission configuration (experiments/submission/                    */
/* pointer-doubling.suite.yaml): local preprocessing on (:471-515,:517-543),*/
/* owner caching on (:666-673, ExtendedReply :37-41), no aggregation        */
/* (DoublingWithoutAggregation :171-250, request_reply_without_remote_      */
/* aggregation request_aggregation_scheme.hpp:156-179), termination by a    */
/* global active count (:410-416).  Block distribution distribution.hpp.    */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t succ;
  int64_t rank;
  int64_t owner;
} pd_reply;

typedef struct {
  int p;
  uint64_t n;
  uint64_t *succ;
  int64_t *rank;
  const uint64_t *off; /* p+1 block offsets */
  pthread_barrier_t bar;
  /* req[src*p+dst]: requested global ids, rep[src*p+dst]: replies in order */
  uint64_t **req;
  uint64_t *req_len;
  uint64_t *req_cap;
  pd_reply **rep;
  uint64_t *active_count; /* per rank */
  int rounds;
} pd_shared;

typedef struct {
  pd_shared *sh;
  int r;
} pd_arg;

static int owner_of(const uint64_t *off, int p, uint64_t g) {
  int lo = 0, hi = p; /* upper_bound(offsets, g) - 1 */
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (off[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

static void push_req(pd_shared *sh, int src, int dst, uint64_t g) {
  int k = src * sh->p + dst;
  if (sh->req_len[k] == sh->req_cap[k]) {
    sh->req_cap[k] = sh->req_cap[k] ? 2 * sh->req_cap[k] : 64;
    sh->req[k] = (uint64_t *)realloc(sh->req[k], sh->req_cap[k] * sizeof(uint64_t));
  }
  sh->req[k][sh->req_len[k]++] = g;
}

typedef struct {
  uint64_t start, next;
  int64_t dist_to_next;
} chain_info;

static void *pd_worker(void *argp) {
  pd_arg *arg = (pd_arg *)argp;
  pd_shared *sh = arg->sh;
  int r = arg->r, p = sh->p;
  uint64_t lo = sh->off[r], hi = sh->off[r + 1], nl = hi - lo;
  uint64_t *succ = sh->succ;
  int64_t *rank = sh->rank;
#define IS_LOCAL(g) ((g) >= lo && (g) < hi)
  /* local_contraction :471-515 */
  unsigned char *has_local_pred = (unsigned char *)calloc(nl ? nl : 1, 1);
  unsigned char *active = (unsigned char *)malloc(nl ? nl : 1);
  memset(active, 1, nl);
  for (uint64_t i = lo; i < hi; i++)
    if (IS_LOCAL(succ[i])) has_local_pred[succ[i] - lo] = 1;
  chain_info *chains = (chain_info *)malloc((nl ? nl : 1) * sizeof(chain_info));
  uint64_t nchains = 0;
  for (uint64_t l = 0; l < nl; l++) {
    if (has_local_pred[l]) continue;
    uint64_t start = lo + l;
    uint64_t cur = succ[start];
    int64_t len = rank[start];
    /* is_end_of_local_chain: non-local, or a root (:491-497) */
    while (IS_LOCAL(cur) && succ[cur] != cur) {
      len += rank[cur];
      active[cur - lo] = 0;
      cur = succ[cur];
    }
    if (cur == succ[start]) continue;
    chains[nchains].start = start;
    chains[nchains].next = succ[start];
    chains[nchains].dist_to_next = rank[start];
    nchains++;
    succ[start] = cur;
    rank[start] = len;
  }
  /* initialize_active_vertices :418-438 */
  uint64_t *act = (uint64_t *)malloc((nl ? nl : 1) * sizeof(uint64_t));
  uint64_t nact = 0;
  for (uint64_t l = 0; l < nl; l++) {
    if (!active[l]) continue;
    uint64_t g = lo + l;
    if (succ[g] == g) succ[g] |= ROOT_FLAG;
    else act[nact++] = g;
  }
  /* owner cache :666-673 */
  int *owner = (int *)malloc((nl ? nl : 1) * sizeof(int));
  for (uint64_t l = 0; l < nl; l++) owner[l] = owner_of(sh->off, p, succ[lo + l] & ~ROOT_FLAG);
  pthread_barrier_wait(&sh->bar);
  int rounds = 0;
  for (;;) {
    /* is_finished :410-416 (allreduce of the active count) */
    sh->active_count[r] = nact;
    pthread_barrier_wait(&sh->bar);
    uint64_t total = 0;
    for (int q = 0; q < p; q++) total += sh->active_count[q];
    pthread_barrier_wait(&sh->bar);
    if (total == 0) break;
    /* make_requests_with_target :122-132 (cached owners) */
    for (int q = 0; q < p; q++) sh->req_len[r * p + q] = 0;
    for (uint64_t k = 0; k < nact; k++) {
      uint64_t g = act[k];
      push_req(sh, r, owner[g - lo], succ[g]);
    }
    pthread_barrier_wait(&sh->bar); /* alltoallv of requests */
    /* make_reply :220-225 -- built before any rank updates (Jacobi) */
    for (int src = 0; src < p; src++) {
      int k = src * p + r;
      sh->rep[k] = (pd_reply *)realloc(sh->rep[k], (sh->req_len[k] ? sh->req_len[k] : 1) * sizeof(pd_reply));
      for (uint64_t j = 0; j < sh->req_len[k]; j++) {
        uint64_t g = sh->req[k][j];
        sh->rep[k][j].succ = succ[g];
        sh->rep[k][j].rank = rank[g];
        sh->rep[k][j].owner = owner[g - lo];
      }
    }
    pthread_barrier_wait(&sh->bar); /* alltoallv of replies */
    /* update_via_read_offsets :62-81 */
    uint64_t read_off[1024];
    for (int q = 0; q < p; q++) read_off[q] = 0;
    uint64_t keep = 0;
    for (uint64_t k = 0; k < nact; k++) {
      uint64_t g = act[k];
      int t = owner[g - lo];
      pd_reply rp = sh->rep[r * p + t][read_off[t]++];
      succ[g] = rp.succ;
      rank[g] += rp.rank;
      owner[g - lo] = (int)rp.owner;
      if (!(succ[g] & ROOT_FLAG)) act[keep++] = g;
    }
    nact = keep;
    rounds++;
    pthread_barrier_wait(&sh->bar);
  }
  /* clear result :690-692 */
  for (uint64_t g = lo; g < hi; g++) succ[g] &= ~ROOT_FLAG;
  /* local_uncontraction :517-543 (only touches local chain interiors) */
  for (uint64_t c = 0; c < nchains; c++) {
    uint64_t start = chains[c].start;
    uint64_t cur = chains[c].next;
    int64_t cur_dist = rank[start] - chains[c].dist_to_next;
    while (IS_LOCAL(cur) && succ[cur] != cur) {
      uint64_t next_node = succ[cur];
      int64_t next_dist = rank[cur];
      succ[cur] = succ[start];
      rank[cur] = cur_dist;
      cur = next_node;
      cur_dist -= next_dist;
    }
  }
#undef IS_LOCAL
  if (r == 0) sh->rounds = rounds;
  free(has_local_pred);
  free(active);
  free(chains);
  free(act);
  free(owner);
  return NULL;
}

/* In place: succ -> root, rank (weights, 0 on roots) -> rank.  Returns the
 * number of doubling rounds, or -1 on bad arguments. */
int orc_pointer_doubling(uint64_t *succ, int64_t *rank, uint64_t n, int p) {
  if (p < 1 || p > 1024) return -1;
  if ((uint64_t)p > n) p = n ? (int)n : 1;
  uint64_t *off = (uint64_t *)malloc((p + 1) * sizeof(uint64_t));
  for (int r = 0; r <= p; r++) off[r] = (uint64_t)r * (n / p) + ((uint64_t)r < n % p ? (uint64_t)r : n % p);
  pd_shared sh;
  memset(&sh, 0, sizeof(sh));
  sh.p = p;
  sh.n = n;
  sh.succ = succ;
  sh.rank = rank;
  sh.off = off;
  pthread_barrier_init(&sh.bar, NULL, (unsigned)p);
  sh.req = (uint64_t **)calloc((size_t)p * p, sizeof(uint64_t *));
  sh.req_len = (uint64_t *)calloc((size_t)p * p, sizeof(uint64_t));
  sh.req_cap = (uint64_t *)calloc((size_t)p * p, sizeof(uint64_t));
  sh.rep = (pd_reply **)calloc((size_t)p * p, sizeof(pd_reply *));
  sh.active_count = (uint64_t *)calloc((size_t)p, sizeof(uint64_t));
  pthread_t *th = (pthread_t *)malloc(p * sizeof(pthread_t));
  pd_arg *args = (pd_arg *)malloc(p * sizeof(pd_arg));
  for (int r = 0; r < p; r++) {
    args[r].sh = &sh;
    args[r].r = r;
    pthread_create(&th[r], NULL, pd_worker, &args[r]);
  }
  for (int r = 0; r < p; r++) pthread_join(th[r], NULL);
  for (int k = 0; k < p * p; k++) {
    free(sh.req[k]);
    free(sh.rep[k]);
  }
  free(sh.req);
  free(sh.req_len);
  free(sh.req_cap);
  free(sh.rep);
  free(sh.active_count);
  free(th);
  free(args);
  free(off);
  pthread_barrier_destroy(&sh.bar);
  return sh.rounds;
}

/* ------------------------------------------------------------------------ */
/* pack / unpack on one address space -- lib/kascade/pack.hpp:23-177.       */
/* `active` lists the active global ids in ascending (rank-major) order; the */
/* packed id of active[k] is k.  Every active element's successor must be    */
/* active (as in the sparse ruling set's ruler sublist).                     */
/* ------------------------------------------------------------------------ */
static int64_t find_sorted(const uint64_t *a, uint64_t len, uint64_t key) {
  uint64_t lo = 0, hi = len;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < len && a[lo] == key) ? (int64_t)lo : -1;
}
int orc_pack(const uint64_t *succ, const int64_t *rank, const uint64_t *active, uint64_t nactive,
             uint64_t *psucc, int64_t *prank) {
  for (uint64_t k = 0; k < nactive; k++) { /* :37-43 */
    psucc[k] = succ[active[k]];
    prank[k] = rank[active[k]];
  }
  for (uint64_t k = 0; k < nactive; k++) { /* :71-80 translate successors */
    int64_t pk = find_sorted(active, nactive, psucc[k]);
    if (pk < 0) return -1;
    psucc[k] = (uint64_t)pk;
  }
  return 0;
}
void orc_unpack(const uint64_t *psucc, const int64_t *prank, const uint64_t *active, uint64_t nactive,
                uint64_t *succ, int64_t *rank) { /* :157-173 write_unpacked */
  for (uint64_t k = 0; k < nactive; k++) {
    succ[active[k]] = active[psucc[k]];
    rank[active[k]] = prank[k];
  }
}

/* ------------------------------------------------------------------------ */
/* Euler-tour tree rooting on one rank -- lib/kascade/eulertour.cpp.         */
/*   reverse_rooted_tree(add_back_edge=true) successor_utils.cpp:265-300,    */
/*   make_graph CSR :236-260, compute_euler_tour_impl eulertour.cpp:33-127,  */
/*   GatherChase ranking (local_pointer_chasing), map_euler_tour_back        */
/*   :195-278.  In place: succ (parent array, roots self-parent) -> root;    */
/*   rank -> depth.  Input weights are ignored (:45-53).                    */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t src, dst, pos;
} et_edge;
static int et_cmp(const void *a, const void *b) {
  const et_edge *x = (const et_edge *)a, *y = (const et_edge *)b;
  if (x->src != y->src) return x->src < y->src ? -1 : 1;
  if (x->dst != y->dst) return x->dst < y->dst ? -1 : 1;
  return 0;
}
/* The reference's tour of a parent array on one PE (compute_euler_tour,
 * eulertour.cpp:379-392, via reverse_rooted_tree successor_utils.cpp:265-300,
 * make_graph :236-260 and compute_euler_tour_impl eulertour.cpp:33-127):
 * one element per directed edge in CSR order (edge id = CSR position), then
 * one dummy per root in vertex order (roots without children keep their
 * self-looping slot at the end).  Fills tsucc/tw/up/efirst/esecond (ne =
 * 2(n - #roots) + #roots entries, allocated here); returns -1 on failure. */
static int reference_tour(const uint64_t *succ, uint64_t n, uint64_t **tsucc_o, int64_t **tw_o,
                          unsigned char **up_o, uint64_t **efirst_o, uint64_t **esecond_o, uint64_t *ne_o) {
  uint64_t nroots = 0, m = 0;
  for (uint64_t i = 0; i < n; i++)
    if (succ[i] == i) nroots++;
  m = 2 * (n - nroots); /* directed edges */
  /* reverse_rooted_tree: per non-root i, back edge (i, p) then (p, i) */
  uint64_t *esrc = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  uint64_t *edst = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  uint64_t e = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (succ[i] == i) continue;
    esrc[e] = i; edst[e] = succ[i]; e++;
    esrc[e] = succ[i]; edst[e] = i; e++;
  }
  /* make_graph: counting sort by src, filling each bucket back to front */
  uint64_t *xadj = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  uint64_t *adj = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  for (uint64_t k = 0; k < m; k++) xadj[esrc[k]]++;
  for (uint64_t v = 1; v <= n; v++) xadj[v] += xadj[v - 1]; /* inclusive scan */
  for (uint64_t k = 0; k < m; k++) adj[--xadj[esrc[k]]] = edst[k];
  /* now xadj[v] is the bucket start, xadj[n] == m */
  /* edge id of (v, adj[pos]) is pos; lookup table for (src, dst) -> pos */
  et_edge *tab = (et_edge *)malloc((m ? m : 1) * sizeof(et_edge));
  for (uint64_t v = 0; v < n; v++)
    for (uint64_t pos = xadj[v]; pos < xadj[v + 1]; pos++) {
      tab[pos].src = v; tab[pos].dst = adj[pos]; tab[pos].pos = pos;
    }
  qsort(tab, m, sizeof(et_edge), et_cmp);
  uint64_t ne = m + nroots;
  uint64_t *tsucc = (uint64_t *)malloc((ne ? ne : 1) * sizeof(uint64_t));
  int64_t *tw = (int64_t *)calloc(ne ? ne : 1, sizeof(int64_t));
  unsigned char *up = (unsigned char *)calloc(ne ? ne : 1, 1);
  uint64_t *efirst = (uint64_t *)malloc((ne ? ne : 1) * sizeof(uint64_t));
  uint64_t *esecond = (uint64_t *)malloc((ne ? ne : 1) * sizeof(uint64_t));
  for (uint64_t k = 0; k < ne; k++) tsucc[k] = k; /* iota :116 */
  uint64_t root_index = m;
  int ok = 1;
  for (uint64_t v = 0; v < n && ok; v++) { /* compute_euler_tour_impl :76-111 */
    uint64_t deg = xadj[v + 1] - xadj[v];
    int is_v_root = (succ[v] == v);
    for (uint64_t i = 0; i < deg; i++) {
      uint64_t u_prev = adj[xadj[v] + (i + deg - 1) % deg];
      uint64_t u = adj[xadj[v] + i];
      uint64_t id = xadj[v] + i;
      efirst[id] = v; esecond[id] = u;
      up[id] = (succ[v] == u);
      tw[id] = up[id] ? 1 : -1; /* assign_weight :45-53, no proxies */
      et_edge key = {u_prev, v, 0};
      et_edge *hit = (et_edge *)bsearch(&key, tab, m, sizeof(et_edge), et_cmp);
      if (!hit) { ok = 0; break; }
      if (is_v_root && i == 0) { /* :92-104 break the tour at the root */
        tsucc[hit->pos] = root_index;
        tw[root_index] = 0;
        up[root_index] = 0;
        efirst[root_index] = v; esecond[root_index] = v;
        root_index++;
      } else {
        tsucc[hit->pos] = id;
      }
    }
  }
  free(esrc); free(edst); free(xadj); free(adj); free(tab);
  if (!ok) {
    free(tsucc); free(tw); free(up); free(efirst); free(esecond);
    return -1;
  }
  *tsucc_o = tsucc; *tw_o = tw; *up_o = up; *efirst_o = efirst; *esecond_o = esecond; *ne_o = ne;
  return 0;
}

/* compute_euler_tour (eulertour.cpp:379-392): the tour successor array (and
 * the +1 / -1 / 0 weights compute_euler_tour_impl assigns, :45-53) into
 * arrays of capacity 2n; returns the tour's length, or -1. */
int64_t orc_compute_euler_tour(const uint64_t *parent, uint64_t n, uint64_t *tsucc_out, int64_t *tw_out) {
  uint64_t *tsucc, *efirst, *esecond, ne;
  int64_t *tw;
  unsigned char *up;
  if (reference_tour(parent, n, &tsucc, &tw, &up, &efirst, &esecond, &ne) != 0) return -1;
  memcpy(tsucc_out, tsucc, ne * sizeof(uint64_t));
  if (tw_out) memcpy(tw_out, tw, ne * sizeof(int64_t));
  free(tsucc); free(tw); free(up); free(efirst); free(esecond);
  return (int64_t)ne;
}

int orc_euler_tour_rooting(uint64_t *succ, int64_t *rank, uint64_t n) {
  /* set_initial_ranking_state (list_ranking.cpp:60-80) */
  for (uint64_t i = 0; i < n; i++) rank[i] = (succ[i] == i) ? 0 : 1;
  uint64_t *tsucc, *efirst, *esecond, ne;
  int64_t *tw;
  unsigned char *up;
  if (reference_tour(succ, n, &tsucc, &tw, &up, &efirst, &esecond, &ne) != 0) return -1;
  orc_local_pointer_chasing(tsucc, tw, ne);
  /* map_euler_tour_back :235-278: upward edge (v, parent) -> (root, depth) */
  for (uint64_t k = 0; k < ne; k++) {
    if (!up[k]) continue;
    uint64_t v = efirst[k];
    succ[v] = esecond[tsucc[k]]; /* request_original_succesors: dummy -> root */
    rank[v] = tw[k];
  }
  free(tsucc); free(tw); free(up); free(efirst); free(esecond);
  return 0;
}
