/*
 * eulertour_mt.c -- CPU restatement of KAscade's rank_via_euler_tour
 * (lib/kascade/eulertour.cpp:337-377) on p threads, for the CPU baseline of
 * tree rooting.  TEST INFRASTRUCTURE ONLY, like oracle.c: a checker and a
 * timed baseline, never part of the kascade_amd product.
 *
 * The reference builds the tour on p MPI ranks (reverse_rooted_tree and
 * make_graph, successor_utils.cpp:236-300; compute_euler_tour_impl,
 * eulertour.cpp:33-127) and ranks it with the sparse ruling set in the
 * submission suite (experiments/eulertour-sparse-ruling-set.suite.yaml:86-92).
 * Here the build runs on p threads over one address space, block by block:
 *   1. degrees: one slot for the parent edge plus one per child (atomic
 *      counts), then a blocked exclusive scan into the CSR offsets;
 *   2. adjacency: slot 0 of every non-root vertex is its parent, children
 *      are appended after it in whatever order the atomic counters give
 *      (any cyclic order of a vertex's neighbours is a valid tour, and the
 *      roots and depths do not depend on it);
 *   3. linking (eulertour.cpp:76-111): the edge (u_prev, v) continues with
 *      (v, u) where u_prev precedes u in v's cyclic order.  The id of the
 *      edge (u_prev, v) needs no hash map: it is u_prev's slot 0 when u_prev
 *      is a child of v, and the slot where v was appended to its parent's
 *      list otherwise.  At a root the edge into its first neighbour is cut
 *      and sent to a per-root dummy element (:92-104);  weights +1 on upward
 *      edges, -1 on downward ones (:45-53);
 *   4. ranking: orc_sparse_ruling_set on p emulated ranks (srs.c);
 *   5. map back (:235-278): the upward edge of v carries depth(v), and its
 *      tour root is the dummy of v's tree root.
 * In place: succ (parent array, roots self-parent) -> root; rank -> depth.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_sparse_ruling_set(uint64_t *succ, int64_t *rank, uint64_t n, int p, int levels);

typedef struct {
  int p;
  uint64_t n, m, ne;
  const uint64_t *par;
  uint64_t *cnt, *fill, *xadj, *adj, *pos_in_parent;
  uint64_t *tsucc, *dummy_vertex, *block_sum, *block_roots;
  int64_t *tw;
  int bad;
  pthread_barrier_t bar;
} et_world;

typedef struct {
  et_world *W;
  int r;
} et_arg;

static void et_block(const et_world *W, int r, uint64_t *lo, uint64_t *hi) {
  const uint64_t per = (W->n + (uint64_t)W->p - 1) / (uint64_t)W->p;
  *lo = per * (uint64_t)r < W->n ? per * (uint64_t)r : W->n;
  *hi = *lo + per < W->n ? *lo + per : W->n;
}

static void *et_build(void *vp) {
  et_arg *a = (et_arg *)vp;
  et_world *W = a->W;
  const int r = a->r;
  uint64_t lo, hi;
  et_block(W, r, &lo, &hi);
  const uint64_t *par = W->par;
  /* 1. degrees */
  uint64_t roots = 0;
  for (uint64_t v = lo; v < hi; v++) {
    const uint64_t p = par[v];
    if (p >= W->n) {
      __atomic_store_n(&W->bad, 1, __ATOMIC_RELAXED);
      continue;
    }
    if (p == v) {
      roots++;
      continue;
    }
    __atomic_fetch_add(&W->cnt[v], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&W->cnt[p], 1, __ATOMIC_RELAXED);
  }
  pthread_barrier_wait(&W->bar);
  uint64_t sum = 0;
  for (uint64_t v = lo; v < hi; v++) sum += W->cnt[v];
  W->block_sum[r] = sum;
  W->block_roots[r] = roots;
  pthread_barrier_wait(&W->bar);
  if (r == 0) {
    uint64_t s = 0, q = 0;
    for (int k = 0; k < W->p; k++) {
      const uint64_t a1 = W->block_sum[k], a2 = W->block_roots[k];
      W->block_sum[k] = s;
      W->block_roots[k] = q;
      s += a1;
      q += a2;
    }
    W->m = s;
    W->ne = s + q;
    W->xadj[W->n] = s;
  }
  pthread_barrier_wait(&W->bar);
  uint64_t off = W->block_sum[r];
  for (uint64_t v = lo; v < hi; v++) {
    W->xadj[v] = off;
    off += W->cnt[v];
  }
  pthread_barrier_wait(&W->bar);
  if (W->bad) return NULL;
  /* 2. adjacency: parent in slot 0, children after it */
  for (uint64_t v = lo; v < hi; v++) {
    const uint64_t p = par[v];
    if (p == v) continue;
    W->adj[W->xadj[v]] = p;
    const uint64_t first = W->xadj[p] + (par[p] != p ? 1 : 0);
    const uint64_t pos = first + __atomic_fetch_add(&W->fill[p], 1, __ATOMIC_RELAXED);
    W->adj[pos] = v;
    W->pos_in_parent[v] = pos;
  }
  pthread_barrier_wait(&W->bar);
  /* 3. linking */
  uint64_t dummy = W->m + W->block_roots[r];
  for (uint64_t v = lo; v < hi; v++) {
    const uint64_t base = W->xadj[v], deg = W->xadj[v + 1] - base;
    const int is_root = par[v] == v;
    uint64_t d = 0;
    if (is_root) {
      d = dummy++;
      W->tsucc[d] = d;
      W->tw[d] = 0;
      W->dummy_vertex[d - W->m] = v;
    }
    for (uint64_t i = 0; i < deg; i++) {
      const uint64_t id = base + i;
      const uint64_t up_prev = W->adj[base + (i + deg - 1) % deg];
      /* id of the edge (u_prev, v) */
      const uint64_t into = (!is_root && up_prev == par[v]) ? W->pos_in_parent[v] : W->xadj[up_prev];
      W->tsucc[into] = (is_root && i == 0) ? d : id;
      W->tw[id] = (!is_root && i == 0) ? 1 : -1;
    }
  }
  return NULL;
}

int orc_euler_tour_rooting_mt(uint64_t *succ, int64_t *rank, uint64_t n, int p) {
  if (p < 1) p = 1;
  if (n == 0) return 0;
  et_world W;
  memset(&W, 0, sizeof W);
  W.p = p;
  W.n = n;
  W.par = succ;
  W.cnt = (uint64_t *)calloc(n, sizeof(uint64_t));
  W.fill = (uint64_t *)calloc(n, sizeof(uint64_t));
  W.xadj = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  W.adj = (uint64_t *)malloc(2 * n * sizeof(uint64_t));
  W.pos_in_parent = (uint64_t *)malloc(n * sizeof(uint64_t));
  W.tsucc = (uint64_t *)malloc(2 * n * sizeof(uint64_t));
  W.tw = (int64_t *)malloc(2 * n * sizeof(int64_t));
  W.dummy_vertex = (uint64_t *)malloc(n * sizeof(uint64_t));
  W.block_sum = (uint64_t *)malloc((size_t)p * sizeof(uint64_t));
  W.block_roots = (uint64_t *)malloc((size_t)p * sizeof(uint64_t));
  int rc = 0;
  pthread_barrier_init(&W.bar, NULL, (unsigned)p);
  pthread_t *th = (pthread_t *)malloc((size_t)p * sizeof(pthread_t));
  et_arg *args = (et_arg *)malloc((size_t)p * sizeof(et_arg));
  for (int r = 0; r < p; r++) {
    args[r].W = &W;
    args[r].r = r;
    pthread_create(&th[r], NULL, et_build, &args[r]);
  }
  for (int r = 0; r < p; r++) pthread_join(th[r], NULL);
  pthread_barrier_destroy(&W.bar);
  if (W.bad) {
    rc = -1;
    goto out;
  }
  /* 4. ranking */
  if (orc_sparse_ruling_set(W.tsucc, W.tw, W.ne, p, 2) != 0) {
    rc = -1;
    goto out;
  }
  /* 5. map back: slot 0 of a non-root vertex is its upward edge */
  for (uint64_t v = 0; v < n; v++) {
    if (succ[v] == v) {
      rank[v] = 0;
      continue;
    }
    const uint64_t up = W.xadj[v];
    rank[v] = W.tw[up];
    succ[v] = W.dummy_vertex[W.tsucc[up] - W.m];
  }
out:
  free(W.cnt); free(W.fill); free(W.xadj); free(W.adj); free(W.pos_in_parent);
  free(W.tsucc); free(W.tw); free(W.dummy_vertex); free(W.block_sum); free(W.block_roots);
  free(th); free(args);
  return rc;
}
