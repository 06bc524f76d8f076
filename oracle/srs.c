/*
 * srs.c -- CPU restatement of KAscade's distributed sparse ruling set on p
 * emulated MPI ranks (one thread per rank, message exchange through shared
 * buffers and barriers).  TEST INFRASTRUCTURE ONLY, like oracle.c: the CPU
 * baseline and a checker, never part of the kascade_amd product.
 *
 * Restates the submission configuration
 * (experiments/submission/sparse-ruling-set-indirection.suite.yaml:60-74):
 * ruler selection ultimate-sanders f = 2.0, spawn, sync engine with
 * locality-aware local forwarding, post-invert, cache owners (a routing
 * detail: results do not depend on it), local contraction, two ruling-set
 * rounds above a pointer-doubling base (the harness's promotion,
 * bench/detail/algorithm_impl.hpp:165-171), pointer-doubling gather fallback
 * ratio 3.
 *
 * Internals that only an import could pin (the libstdc++ mt19937 shuffle that
 * picks rulers, sparse_ruling_set.cpp:106-117) are replaced by a seeded
 * xorshift shuffle: the chosen rulers differ, the outputs (unique for a list
 * forest) do not.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FLAG (UINT64_C(1) << 63)

enum { T_ROOT, T_LEAF, T_RULER, T_UNREACHED, T_REACHED, T_MASKED }; /* types.hpp:10 */

/* ------------------------------------------------------------------ world */
typedef struct {
  int p;
  pthread_barrier_t bar;
  /* msg[src*p+dst]: raw byte buffers */
  unsigned char **msg;
  uint64_t *len, *cap;
  uint64_t *scalar; /* per-rank scalars for collectives */
} world;

static void w_barrier(world *W) { pthread_barrier_wait(&W->bar); }

static uint64_t w_allreduce_sum(world *W, int r, uint64_t x) {
  W->scalar[r] = x;
  w_barrier(W);
  uint64_t s = 0;
  for (int q = 0; q < W->p; q++) s += W->scalar[q];
  w_barrier(W);
  return s;
}

/* allgather of one u64 per rank into out[p] */
static void w_allgather(world *W, int r, uint64_t x, uint64_t *out) {
  W->scalar[r] = x;
  w_barrier(W);
  for (int q = 0; q < W->p; q++) out[q] = W->scalar[q];
  w_barrier(W);
}

static void w_push(world *W, int src, int dst, const void *rec, uint64_t sz) {
  int k = src * W->p + dst;
  if (W->len[k] + sz > W->cap[k]) {
    uint64_t c = W->cap[k] ? W->cap[k] * 2 : 1024;
    while (c < W->len[k] + sz) c *= 2;
    W->msg[k] = (unsigned char *)realloc(W->msg[k], c);
    W->cap[k] = c;
  }
  memcpy(W->msg[k] + W->len[k], rec, sz);
  W->len[k] += sz;
}

/* Alltoallv: after w_exchange_begin, rank r reads every W->msg[src*p+r];
 * w_exchange_end clears r's outgoing buffers for the next exchange. */
static void w_exchange_begin(world *W) { w_barrier(W); }
static void w_exchange_end(world *W, int r) {
  w_barrier(W);
  for (int q = 0; q < W->p; q++) W->len[r * W->p + q] = 0;
  w_barrier(W);
}

static int owner(const uint64_t *off, int p, uint64_t g) {
  int lo = 0, hi = p;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (off[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

/* per-level shared description: block offsets of the level's arrays */
typedef struct {
  uint64_t *succ;
  int64_t *rank;
  uint64_t *off; /* p+1 */
} level;

/* ------------------------------------------------- pointer doubling base */
typedef struct {
  uint64_t succ;
  int64_t rank;
} pd_reply;

/* gather fallback: rank_on_root (list_ranking.cpp:29-58) = sequential walk */
void orc_local_pointer_chasing(uint64_t *succ, int64_t *rank, uint64_t n);

/* pointer_doubling_generic.hpp:608-700 restricted to what the base case of
 * the submission configuration runs: fallback when n < 3p (:624-634), else
 * Jacobi rounds with request/reply.  Local contraction is on in that
 * configuration; it is a per-rank sequential optimisation whose outputs equal
 * the plain rounds', and the rulers' sublist has few PE-local successors. */
static void pd_rank(world *W, int r, level L) {
  int p = W->p;
  uint64_t n = L.off[p], lo = L.off[r], hi = L.off[r + 1];
  if (n < (uint64_t)(3 * p)) {
    w_barrier(W);
    if (r == 0) orc_local_pointer_chasing(L.succ, L.rank, n);
    w_barrier(W);
    return;
  }
  uint64_t nl = hi - lo;
  uint64_t *act = (uint64_t *)malloc((nl ? nl : 1) * sizeof(uint64_t));
  uint64_t nact = 0;
  for (uint64_t g = lo; g < hi; g++) {
    if (L.succ[g] == g) L.succ[g] |= FLAG; /* initialize_active_vertices */
    else act[nact++] = g;
  }
  for (;;) {
    if (w_allreduce_sum(W, r, nact) == 0) break;
    for (uint64_t k = 0; k < nact; k++) {
      uint64_t s = L.succ[act[k]];
      w_push(W, r, owner(L.off, p, s), &s, sizeof s);
    }
    w_exchange_begin(W);
    /* replies from the pre-round state (Jacobi) into a private buffer */
    uint64_t tot = 0;
    for (int src = 0; src < p; src++) tot += W->len[src * p + r] / sizeof(uint64_t);
    pd_reply *rep = (pd_reply *)malloc((tot ? tot : 1) * sizeof(pd_reply));
    uint64_t *rep_off = (uint64_t *)calloc(p + 1, sizeof(uint64_t));
    uint64_t k = 0;
    for (int src = 0; src < p; src++) {
      rep_off[src] = k;
      uint64_t *q = (uint64_t *)W->msg[src * p + r];
      uint64_t cnt = W->len[src * p + r] / sizeof(uint64_t);
      for (uint64_t j = 0; j < cnt; j++, k++) {
        rep[k].succ = L.succ[q[j]];
        rep[k].rank = L.rank[q[j]];
      }
    }
    rep_off[p] = k;
    w_exchange_end(W, r);
    for (int src = 0; src < p; src++)
      for (uint64_t j = rep_off[src]; j < rep_off[src + 1]; j++) w_push(W, r, src, &rep[j], sizeof(pd_reply));
    free(rep);
    free(rep_off);
    w_exchange_begin(W);
    uint64_t *rd = (uint64_t *)calloc(p, sizeof(uint64_t));
    uint64_t keep = 0;
    for (uint64_t a = 0; a < nact; a++) {
      uint64_t g = act[a];
      int t = owner(L.off, p, L.succ[g]);
      pd_reply rp;
      memcpy(&rp, W->msg[t * p + r] + rd[t] * sizeof(pd_reply), sizeof rp);
      rd[t]++;
      L.succ[g] = rp.succ;
      L.rank[g] += rp.rank;
      if (!(rp.succ & FLAG)) act[keep++] = g;
    }
    free(rd);
    nact = keep;
    w_exchange_end(W, r);
  }
  for (uint64_t g = lo; g < hi; g++) L.succ[g] &= ~FLAG;
  free(act);
  w_barrier(W);
}

/* ------------------------------------------------------ sparse ruling set */
typedef struct {
  uint64_t target, ruler;
  int64_t dist;
} ruler_msg; /* RulerMessage, types.hpp:30-34 */

typedef struct {
  uint64_t pred, succ;
  int64_t dist;
} fix_msg; /* post_invert.hpp:216-220 */

typedef struct {
  uint64_t leaf, root;
  int64_t dist;
} leaf_msg; /* LeafMessage, post_invert.hpp:24-28 */

typedef struct {
  uint64_t start, next;
  int64_t dist_to_next;
} chain_info;

static uint64_t xs64(uint64_t *s) {
  uint64_t x = *s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return *s = x;
}

typedef struct {
  world *W;
  int r;
  int levels; /* sparse_ruling_set_rounds */
  double factor;
} srs_ctx;

static void srs_rank(srs_ctx *C, level L, int cur_level);

/* a vector of u64 */
typedef struct {
  uint64_t *a;
  uint64_t n, cap;
} vec;
static void vpush(vec *v, uint64_t x) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 64;
    v->a = (uint64_t *)realloc(v->a, v->cap * sizeof(uint64_t));
  }
  v->a[v->n++] = x;
}

static void srs_rank(srs_ctx *C, level L, int cur_level) {
  world *W = C->W;
  int r = C->r, p = W->p;
  uint64_t lo = L.off[r], hi = L.off[r + 1], nl = hi - lo;
  uint64_t *succ = L.succ;
  int64_t *rank = L.rank;
#define IS_LOCAL(g) ((g) >= lo && (g) < hi)

  /* node types (sparse_ruling_set.cpp:67-86); post-invert without leaf detection */
  unsigned char *type = (unsigned char *)malloc(nl ? nl : 1);
  uint64_t num_unreached = 0;
  for (uint64_t l = 0; l < nl; l++) {
    type[l] = succ[lo + l] == lo + l ? T_ROOT : T_UNREACHED;
    if (type[l] == T_UNREACHED) num_unreached++;
  }

  /* local contraction (local_contraction.hpp:28-75) */
  unsigned char *has_local_pred = (unsigned char *)calloc(nl ? nl : 1, 1);
  for (uint64_t l = 0; l < nl; l++)
    if (IS_LOCAL(succ[lo + l])) has_local_pred[succ[lo + l] - lo] = 1;
  chain_info *chains = (chain_info *)malloc((nl ? nl : 1) * sizeof(chain_info));
  uint64_t nchains = 0, num_masked = 0;
  for (uint64_t l = 0; l < nl; l++) {
    if (has_local_pred[l]) continue;
    uint64_t start = lo + l, cur = succ[start];
    int64_t len = rank[start];
    while (IS_LOCAL(cur) && type[cur - lo] != T_ROOT) {
      len += rank[cur];
      type[cur - lo] = T_MASKED;
      num_masked++;
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
  free(has_local_pred);
  num_unreached -= num_masked;

  /* permutation of the local unreached indices (replaces mt19937 + shuffle) */
  uint64_t *perm = (uint64_t *)malloc((nl ? nl : 1) * sizeof(uint64_t));
  uint64_t nperm = 0;
  for (uint64_t l = 0; l < nl; l++)
    if (type[l] == T_UNREACHED) perm[nperm++] = l;
  uint64_t seed = 0x9E3779B97F4A7C15ull ^ (uint64_t)(42 + r) ^ ((uint64_t)cur_level << 40);
  for (uint64_t i = nperm; i > 1; i--) {
    uint64_t j = xs64(&seed) % i;
    uint64_t t = perm[i - 1];
    perm[i - 1] = perm[j];
    perm[j] = t;
  }

  /* compute_local_num_rulers, ultimate_sanders with d = 1 (ruler_selection.hpp:55-63) */
  uint64_t n_local = nl - num_masked;
  uint64_t n = w_allreduce_sum(W, r, n_local);
  double rel = n ? (double)n_local / (double)n : 0.0;
  double rtot = p > 1 ? C->factor * sqrt((double)n * pow((double)p, 2.0)) / log2((double)p) : sqrt((double)n);
  uint64_t k_rulers = (uint64_t)(rtot * rel);

  /* pick_rulers (:81-115): the first k unreached of the permutation */
  vec rulers = {0};
  uint64_t next_perm = 0;
  for (; next_perm < nperm && rulers.n < k_rulers; next_perm++)
    if (type[perm[next_perm]] == T_UNREACHED) vpush(&rulers, perm[next_perm]);
  num_unreached -= rulers.n;

  /* ruler chase, sync engine (ruler_chasing_engine.hpp:57-101) with
   * locality-aware local forwarding and spawn (sparse_ruling_set.cpp:157-265) */
  ruler_msg *local_work = NULL;
  uint64_t lw_n = 0, lw_cap = 0;
#define LOCAL_PUSH(m)                                                               \
  do {                                                                              \
    if (lw_n == lw_cap) {                                                           \
      lw_cap = lw_cap ? 2 * lw_cap : 256;                                           \
      local_work = (ruler_msg *)realloc(local_work, lw_cap * sizeof(ruler_msg));   \
    }                                                                               \
    local_work[lw_n++] = (m);                                                       \
  } while (0)
  uint64_t outgoing = 0;
  for (uint64_t i = 0; i < rulers.n; i++) { /* init: everything through the buffers */
    uint64_t l = rulers.a[i], g = lo + l;
    type[l] = T_RULER;
    ruler_msg m = {succ[g], g, rank[g]};
    succ[g] = g;
    rank[g] = 0;
    w_push(W, r, owner(L.off, p, m.target), &m, sizeof m);
    outgoing++;
  }
  for (;;) {
    if (w_allreduce_sum(W, r, outgoing) == 0) break;
    w_exchange_begin(W);
    lw_n = 0;
    for (int src = 0; src < p; src++) {
      ruler_msg *q = (ruler_msg *)W->msg[src * p + r];
      uint64_t cnt = W->len[src * p + r] / sizeof(ruler_msg);
      for (uint64_t j = 0; j < cnt; j++) LOCAL_PUSH(q[j]);
    }
    w_exchange_end(W, r);
    outgoing = 0;
    for (uint64_t i = 0; i < lw_n; i++) {
      ruler_msg m = local_work[i];
      uint64_t l = m.target - lo;
      uint64_t s = succ[m.target];
      int64_t d = rank[m.target];
      succ[m.target] = m.ruler;
      rank[m.target] = m.dist;
      if (type[l] == T_RULER || type[l] == T_ROOT) {
        /* chain end; spawn a new ruler from the permutation */
        while (next_perm < nperm && type[perm[next_perm]] != T_UNREACHED) next_perm++;
        if (num_unreached > 0 && next_perm < nperm) {
          uint64_t rl = perm[next_perm++], rg = lo + rl;
          vpush(&rulers, rl);
          type[rl] = T_RULER;
          num_unreached--;
          ruler_msg nm = {succ[rg], rg, rank[rg]};
          succ[rg] = rg;
          rank[rg] = 0;
          if (IS_LOCAL(nm.target)) LOCAL_PUSH(nm);
          else {
            w_push(W, r, owner(L.off, p, nm.target), &nm, sizeof nm);
            outgoing++;
          }
        }
        continue;
      }
      type[l] = T_REACHED;
      num_unreached--;
      ruler_msg nm = {s, m.ruler, m.dist + d};
      if (IS_LOCAL(s)) LOCAL_PUSH(nm);
      else {
        w_push(W, r, owner(L.off, p, s), &nm, sizeof nm);
        outgoing++;
      }
    }
  }
  free(local_work);

  /* fixup_unreached (post_invert.hpp:195-264) */
  vec fix = {0};
  for (uint64_t l = 0; l < nl; l++)
    if (type[l] == T_UNREACHED) vpush(&fix, l);
  for (uint64_t i = 0; i < fix.n; i++) {
    uint64_t l = fix.a[i], g = lo + l;
    type[l] = T_LEAF;
    fix_msg m = {g, succ[g], rank[g]};
    succ[g] = g;
    rank[g] = 0;
    w_push(W, r, owner(L.off, p, m.succ), &m, sizeof m);
  }
  w_exchange_begin(W);
  vec extra = {0};
  for (int src = 0; src < p; src++) {
    fix_msg *q = (fix_msg *)W->msg[src * p + r];
    uint64_t cnt = W->len[src * p + r] / sizeof(fix_msg);
    for (uint64_t j = 0; j < cnt; j++) {
      uint64_t l = q[j].succ - lo;
      succ[q[j].succ] = q[j].pred;
      rank[q[j].succ] = q[j].dist;
      if (type[l] == T_ROOT) vpush(&extra, l);
      if (type[l] == T_LEAF) type[l] = T_RULER;
    }
  }
  w_exchange_end(W, r);
  for (uint64_t i = 0; i < fix.n; i++) vpush(&rulers, fix.a[i]);
  for (uint64_t i = 0; i < extra.n; i++) vpush(&rulers, extra.a[i]);
  free(fix.a);
  free(extra.a);

  /* pack (pack.hpp:23-80): packed id of rulers[i] on rank r = poff[r] + i */
  uint64_t *cnts = (uint64_t *)malloc(p * sizeof(uint64_t));
  w_allgather(W, r, rulers.n, cnts);
  uint64_t *poff = (uint64_t *)malloc((p + 1) * sizeof(uint64_t));
  poff[0] = 0;
  for (int q = 0; q < p; q++) poff[q + 1] = poff[q] + cnts[q];
  free(cnts);
  /* shared packed arrays: rank 0 allocates, everyone fills its block */
  static uint64_t *psucc_shared;
  static int64_t *prank_shared;
  static uint64_t *poff_shared;
  w_barrier(W);
  if (r == 0) {
    psucc_shared = (uint64_t *)malloc((poff[p] ? poff[p] : 1) * sizeof(uint64_t));
    prank_shared = (int64_t *)malloc((poff[p] ? poff[p] : 1) * sizeof(int64_t));
    poff_shared = poff;
  }
  w_barrier(W);
  uint64_t *psucc = psucc_shared;
  int64_t *prank = prank_shared;
  uint64_t *pof = poff_shared;
  w_barrier(W);
  if (r != 0) free(poff);
  /* local unpacked -> packed map for owner replies */
  uint64_t *to_packed = (uint64_t *)malloc((nl ? nl : 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < rulers.n; i++) to_packed[rulers.a[i]] = pof[r] + i;
  for (uint64_t i = 0; i < rulers.n; i++) { /* request the packed id of each successor */
    uint64_t s = succ[lo + rulers.a[i]];
    w_push(W, r, owner(L.off, p, s), &s, sizeof s);
  }
  w_exchange_begin(W);
  vec rep[1024];
  for (int q = 0; q < p; q++) rep[q] = (vec){0};
  for (int src = 0; src < p; src++) {
    uint64_t *q = (uint64_t *)W->msg[src * p + r];
    uint64_t cnt = W->len[src * p + r] / sizeof(uint64_t);
    for (uint64_t j = 0; j < cnt; j++) vpush(&rep[src], to_packed[q[j] - lo]);
  }
  w_exchange_end(W, r);
  for (int q = 0; q < p; q++) {
    for (uint64_t j = 0; j < rep[q].n; j++) w_push(W, r, q, &rep[q].a[j], sizeof(uint64_t));
    free(rep[q].a);
  }
  w_exchange_begin(W);
  {
    uint64_t *rd = (uint64_t *)calloc(p, sizeof(uint64_t));
    for (uint64_t i = 0; i < rulers.n; i++) {
      uint64_t g = lo + rulers.a[i];
      int t = owner(L.off, p, succ[g]);
      uint64_t v;
      memcpy(&v, W->msg[t * p + r] + rd[t] * sizeof(uint64_t), sizeof v);
      rd[t]++;
      psucc[pof[r] + i] = v;
      prank[pof[r] + i] = rank[g];
    }
    free(rd);
  }
  w_exchange_end(W, r);

  /* base case (sparse_ruling_set.cpp:300-309, :381-392) */
  level PL = {psucc, prank, pof};
  if (cur_level + 1 < C->levels) srs_rank(C, PL, cur_level + 1);
  else pd_rank(W, r, PL);

  /* unpack (pack.hpp:81-175): packed root ids back to unpacked ids */
  for (uint64_t i = 0; i < rulers.n; i++) {
    uint64_t ps = psucc[pof[r] + i];
    w_push(W, r, owner(pof, p, ps), &ps, sizeof ps);
  }
  w_exchange_begin(W);
  for (int q = 0; q < p; q++) rep[q] = (vec){0};
  for (int src = 0; src < p; src++) {
    uint64_t *q = (uint64_t *)W->msg[src * p + r];
    uint64_t cnt = W->len[src * p + r] / sizeof(uint64_t);
    for (uint64_t j = 0; j < cnt; j++) vpush(&rep[src], lo + rulers.a[q[j] - pof[r]]);
  }
  w_exchange_end(W, r);
  for (int q = 0; q < p; q++) {
    for (uint64_t j = 0; j < rep[q].n; j++) w_push(W, r, q, &rep[q].a[j], sizeof(uint64_t));
    free(rep[q].a);
  }
  w_exchange_begin(W);
  {
    uint64_t *rd = (uint64_t *)calloc(p, sizeof(uint64_t));
    for (uint64_t i = 0; i < rulers.n; i++) {
      uint64_t ps = psucc[pof[r] + i];
      int t = owner(pof, p, ps);
      uint64_t v;
      memcpy(&v, W->msg[t * p + r] + rd[t] * sizeof(uint64_t), sizeof v);
      rd[t]++;
      succ[lo + rulers.a[i]] = v;
      rank[lo + rulers.a[i]] = prank[pof[r] + i];
    }
    free(rd);
  }
  w_exchange_end(W, r);
  w_barrier(W);
  if (r == 0) { /* the level's packed arrays (the statics may now name a deeper level's) */
    free(psucc);
    free(prank);
    free(pof);
  }
  w_barrier(W);
  free(to_packed);

  /* ruler propagation, pull (ruler_propagation.hpp:148-257), locality-aware */
  unsigned char *is_ruler_node = (unsigned char *)calloc(nl ? nl : 1, 1);
  for (uint64_t i = 0; i < rulers.n; i++) is_ruler_node[rulers.a[i]] = 1;
#define NEEDS(l) (!is_ruler_node[l] && type[l] != T_RULER && type[l] != T_LEAF && type[l] != T_MASKED)
  for (uint64_t l = 0; l < nl; l++) {
    if (!NEEDS(l)) continue;
    uint64_t ru = succ[lo + l];
    if (IS_LOCAL(ru)) continue;
    w_push(W, r, owner(L.off, p, ru), &ru, sizeof ru);
  }
  w_exchange_begin(W);
  vec rr[1024];
  for (int q = 0; q < p; q++) rr[q] = (vec){0};
  for (int src = 0; src < p; src++) {
    uint64_t *q = (uint64_t *)W->msg[src * p + r];
    uint64_t cnt = W->len[src * p + r] / sizeof(uint64_t);
    for (uint64_t j = 0; j < cnt; j++) {
      vpush(&rr[src], succ[q[j]]);
      vpush(&rr[src], (uint64_t)rank[q[j]]);
    }
  }
  w_exchange_end(W, r);
  for (int q = 0; q < p; q++) {
    for (uint64_t j = 0; j < rr[q].n; j++) w_push(W, r, q, &rr[q].a[j], sizeof(uint64_t));
    free(rr[q].a);
  }
  w_exchange_begin(W);
  {
    uint64_t *rd = (uint64_t *)calloc(p, sizeof(uint64_t));
    for (uint64_t l = 0; l < nl; l++) {
      if (!NEEDS(l)) continue;
      uint64_t ru = succ[lo + l];
      if (IS_LOCAL(ru)) {
        succ[lo + l] = succ[ru];
        rank[lo + l] += rank[ru];
        continue;
      }
      int t = owner(L.off, p, ru);
      uint64_t v[2];
      memcpy(v, W->msg[t * p + r] + rd[t] * 2 * sizeof(uint64_t), sizeof v);
      rd[t]++;
      succ[lo + l] = v[0];
      rank[lo + l] += (int64_t)v[1];
    }
    free(rd);
  }
  w_exchange_end(W, r);
#undef NEEDS
  free(is_ruler_node);

  /* local uncontraction (local_contraction.hpp:77-101), prefix orientation */
  for (uint64_t c = 0; c < nchains; c++) {
    uint64_t start = chains[c].start, cur = chains[c].next;
    int64_t cur_dist = rank[start] + chains[c].dist_to_next;
    while (IS_LOCAL(cur) && type[cur - lo] != T_ROOT) {
      uint64_t nx = succ[cur];
      int64_t nd = rank[cur];
      succ[cur] = succ[start];
      rank[cur] = cur_dist;
      type[cur - lo] = T_REACHED;
      cur = nx;
      cur_dist += nd;
    }
  }
  free(chains);

  /* post_invert (post_invert.hpp:138-193): each root sends (root, total) to
   * its head; every other node becomes total - prefix */
  for (uint64_t l = 0; l < nl; l++) {
    if (type[l] != T_ROOT) continue;
    uint64_t g = lo + l, leaf = succ[g];
    if (leaf == g) continue;
    leaf_msg m = {leaf, g, rank[g]};
    succ[g] = g;
    rank[g] = 0;
    w_push(W, r, owner(L.off, p, leaf), &m, sizeof m);
  }
  w_exchange_begin(W);
  for (int src = 0; src < p; src++) {
    leaf_msg *q = (leaf_msg *)W->msg[src * p + r];
    uint64_t cnt = W->len[src * p + r] / sizeof(leaf_msg);
    for (uint64_t j = 0; j < cnt; j++) {
      succ[q[j].leaf] = q[j].root;
      rank[q[j].leaf] = q[j].dist;
      type[q[j].leaf - lo] = T_LEAF;
    }
  }
  w_exchange_end(W, r);
  /* leaf info for remote heads: request (root, total) from the head's owner */
  for (uint64_t l = 0; l < nl; l++) {
    if (type[l] == T_ROOT || type[l] == T_LEAF || type[l] == T_MASKED) continue;
    uint64_t leaf = succ[lo + l];
    if (IS_LOCAL(leaf)) continue;
    w_push(W, r, owner(L.off, p, leaf), &leaf, sizeof leaf);
  }
  w_exchange_begin(W);
  for (int q = 0; q < p; q++) rr[q] = (vec){0};
  for (int src = 0; src < p; src++) {
    uint64_t *q = (uint64_t *)W->msg[src * p + r];
    uint64_t cnt = W->len[src * p + r] / sizeof(uint64_t);
    for (uint64_t j = 0; j < cnt; j++) {
      vpush(&rr[src], succ[q[j]]);
      vpush(&rr[src], (uint64_t)rank[q[j]]);
    }
  }
  w_exchange_end(W, r);
  for (int q = 0; q < p; q++) {
    for (uint64_t j = 0; j < rr[q].n; j++) w_push(W, r, q, &rr[q].a[j], sizeof(uint64_t));
    free(rr[q].a);
  }
  w_exchange_begin(W);
  {
    uint64_t *rd = (uint64_t *)calloc(p, sizeof(uint64_t));
    for (uint64_t l = 0; l < nl; l++) {
      if (type[l] == T_ROOT || type[l] == T_LEAF || type[l] == T_MASKED) continue;
      uint64_t leaf = succ[lo + l];
      if (IS_LOCAL(leaf)) {
        succ[lo + l] = succ[leaf];
        rank[lo + l] = rank[leaf] - rank[lo + l];
        continue;
      }
      int t = owner(L.off, p, leaf);
      uint64_t v[2];
      memcpy(v, W->msg[t * p + r] + rd[t] * 2 * sizeof(uint64_t), sizeof v);
      rd[t]++;
      succ[lo + l] = v[0];
      rank[lo + l] = (int64_t)v[1] - rank[lo + l];
    }
    free(rd);
  }
  w_exchange_end(W, r);
  free(type);
  free(perm);
  free(rulers.a);
#undef IS_LOCAL
#undef LOCAL_PUSH
}

typedef struct {
  srs_ctx C;
  level L;
  int pd_only;
} srs_arg;

static void *srs_thread(void *a) {
  srs_arg *A = (srs_arg *)a;
  if (A->pd_only) pd_rank(A->C.W, A->C.r, A->L);
  else srs_rank(&A->C, A->L, 0);
  return NULL;
}

static int run_world(uint64_t *succ, int64_t *rank, uint64_t n, int p, int pd_only, int levels) {
  if (p < 1 || p > 1024) return -1;
  if ((uint64_t)p > n) p = n ? (int)n : 1;
  world W;
  memset(&W, 0, sizeof W);
  W.p = p;
  pthread_barrier_init(&W.bar, NULL, (unsigned)p);
  W.msg = (unsigned char **)calloc((size_t)p * p, sizeof(unsigned char *));
  W.len = (uint64_t *)calloc((size_t)p * p, sizeof(uint64_t));
  W.cap = (uint64_t *)calloc((size_t)p * p, sizeof(uint64_t));
  W.scalar = (uint64_t *)calloc((size_t)p, sizeof(uint64_t));
  uint64_t *off = (uint64_t *)malloc((p + 1) * sizeof(uint64_t));
  for (int r = 0; r <= p; r++) off[r] = (uint64_t)r * (n / p) + ((uint64_t)r < n % p ? (uint64_t)r : n % p);
  pthread_t *th = (pthread_t *)malloc(p * sizeof(pthread_t));
  srs_arg *args = (srs_arg *)malloc(p * sizeof(srs_arg));
  for (int r = 0; r < p; r++) {
    args[r].C.W = &W;
    args[r].C.r = r;
    args[r].C.levels = levels;
    args[r].C.factor = 2.0;
    args[r].L.succ = succ;
    args[r].L.rank = rank;
    args[r].L.off = off;
    args[r].pd_only = pd_only;
    pthread_create(&th[r], NULL, srs_thread, &args[r]);
  }
  for (int r = 0; r < p; r++) pthread_join(th[r], NULL);
  for (int k = 0; k < p * p; k++) free(W.msg[k]);
  free(W.msg);
  free(W.len);
  free(W.cap);
  free(W.scalar);
  free(off);
  free(th);
  free(args);
  pthread_barrier_destroy(&W.bar);
  return 0;
}

/* In place: succ -> root, rank (weights, 0 on roots) -> rank, on p emulated
 * ranks with `levels` ruling-set rounds above the pointer-doubling base. */
int orc_sparse_ruling_set(uint64_t *succ, int64_t *rank, uint64_t n, int p, int levels) {
  if (n == 0) return 0;
  return run_world(succ, rank, n, p, 0, levels < 1 ? 1 : levels);
}
