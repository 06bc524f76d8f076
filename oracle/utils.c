/*
 * utils.c -- CPU restatement of KAscade's successor utilities and of the
 * benchmark statistics.  TEST INFRASTRUCTURE ONLY: used by tests/ as the
 * checker; never linked into, loaded by or called from the kascade_amd
 * product library.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the niklas-uhl/kascade source tree); tests/test_oracle.py pins each one.
 * One address space stands for all MPI ranks: the request / reply exchanges of
 * the reference become direct array accesses, which changes no result.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* is_list, lib/kascade/successor_utils.cpp:31-63: every element that is not a
 * root "requests" its successor; a successor requested twice has two
 * predecessors.  Returns 1 (a list / forest of lists), 0, or -1 for an id >= n. */
int orc_is_list(const uint64_t *succ, uint64_t n) {
  uint8_t *seen = (uint8_t *)calloc(n ? n : 1, 1);
  int ok = 1;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t s = succ[i];
    if (s >= n) { ok = -1; break; }
    if (s == i) continue; /* :41-43 */
    if (seen[s]) ok = 0;  /* :52-58 duplicate request */
    seen[s] = 1;
  }
  free(seen);
  return ok;
}

/* LeafInfo, successor_utils.cpp:72-123: has_pred[succ] for every non-root;
 * is_leaf(v) = !has_pred[v] (:114-117), so a root without predecessor is a
 * leaf too; num_local_leaves counts them.  Returns the count, -1 on a bad id. */
int64_t orc_leaf_info(const uint64_t *succ, uint64_t n, uint8_t *is_leaf) {
  uint8_t *has_pred = (uint8_t *)calloc(n ? n : 1, 1);
  for (uint64_t i = 0; i < n; i++) {
    if (succ[i] >= n) { free(has_pred); return -1; }
    if (succ[i] != i) has_pred[succ[i]] = 1;
  }
  int64_t leaves = 0;
  for (uint64_t v = 0; v < n; v++) {
    is_leaf[v] = !has_pred[v];
    leaves += !has_pred[v];
  }
  free(has_pred);
  return leaves;
}

/* reverse_list, successor_utils.cpp:159-232: every node starts as its own
 * predecessor with distance 0 (:216-218); each non-root u sends (pred = u,
 * succ, dist_to_succ[u]) to its successor, which stores it (:219-230).  The
 * original roots are returned in index order (:188-190).  dist == NULL stands
 * for unit distances.  Returns the number of roots, -1 on a bad id or on a
 * second predecessor (the KASSERT at :183). */
int64_t orc_reverse_list(const uint64_t *succ, const int64_t *dist, uint64_t n, uint64_t *pred,
                         int64_t *dist_to_pred, uint64_t *roots) {
  for (uint64_t v = 0; v < n; v++) {
    pred[v] = v;
    dist_to_pred[v] = 0;
  }
  int64_t nroots = 0;
  for (uint64_t u = 0; u < n; u++) {
    uint64_t s = succ[u];
    if (s >= n) return -1;
    if (s == u) {
      if (roots) roots[nroots] = u;
      nroots++;
      continue;
    }
    if (pred[s] != s) return -1;
    pred[s] = u;
    dist_to_pred[s] = dist ? dist[u] : 1;
  }
  return nroots;
}

/* Layout identical to kc_stats in include/kascade_amd.h. */
typedef struct orc_stats {
  uint64_t num_trees;
  int64_t max_rank;
  double avg_rank;
  uint64_t num_nontrivial_trees;
  uint64_t nontrivial_size_sum;
  int64_t nontrivial_rank_sum;
  uint64_t max_size;
  double avg_size;
  double nontrivial_avg_size;
  double nontrivial_avg_rank;
} orc_stats;

static double safe_division(double a, double b) { return b == 0 ? 0.0 : a / b; } /* stats.cpp:18-22 */

/* compute_stats, bench/detail/stats.cpp:147-162, with level as StatsLevel
 * (none 1, basic 2, extensive-reduced-output 3, extensive 4):
 *   basic (compute_basic_stats :115-145): roots counted by root[i] == i, the
 *     maximum rank and the rank sum / n;
 *   extensive (compute_per_tree_stats :24-113): per root value its size, rank
 *     sum and maximum rank; the aggregates over trees of size > 1.
 * The per-tree list (level 4) is written in ascending root order; the
 * reference's order is its hash map's, which carries no meaning.  Root values
 * must be < n.  Returns the number of per-tree entries, -1 on a bad root. */
int64_t orc_compute_stats(const uint64_t *root, const int64_t *rank, uint64_t n, int level, orc_stats *out,
                          uint64_t *tree_root, uint64_t *tree_size, int64_t *tree_max, int64_t *tree_sum) {
  memset(out, 0, sizeof(*out));
  if (n == 0 || level <= 1) return 0;
  int64_t mx = INT64_MIN, sum = 0;
  uint64_t roots = 0;
  for (uint64_t i = 0; i < n; i++) {
    roots += root[i] == i;
    if (rank[i] > mx) mx = rank[i];
    sum += rank[i];
  }
  out->num_trees = roots;
  out->max_rank = mx;
  out->avg_rank = (double)sum / (double)n;
  if (level == 2) return 0;
  uint64_t *size = (uint64_t *)calloc(n, sizeof(uint64_t));
  int64_t *rsum = (int64_t *)calloc(n, sizeof(int64_t));
  int64_t *rmax = (int64_t *)malloc(n * sizeof(int64_t));
  for (uint64_t i = 0; i < n; i++) rmax[i] = INT64_MIN;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t r = root[i];
    if (r >= n) { free(size); free(rsum); free(rmax); return -1; }
    size[r]++;                               /* :34 */
    rsum[r] += rank[i];                      /* :35 */
    if (rank[i] > rmax[r]) rmax[r] = rank[i]; /* :36 */
  }
  uint64_t trees = 0;
  int64_t k = 0;
  for (uint64_t r = 0; r < n; r++) {
    if (size[r] == 0) continue;
    trees++;
    if (size[r] > out->max_size) out->max_size = size[r];
    if (size[r] > 1) { /* :63-69 */
      out->num_nontrivial_trees++;
      out->nontrivial_size_sum += size[r];
      out->nontrivial_rank_sum += rsum[r];
      if (level == 4) {
        if (tree_root) tree_root[k] = r;
        if (tree_size) tree_size[k] = size[r];
        if (tree_max) tree_max[k] = rmax[r];
        if (tree_sum) tree_sum[k] = rsum[r];
      }
      k++;
    }
  }
  out->num_trees = trees;
  uint64_t trivial = trees - out->num_nontrivial_trees; /* :92-100 */
  out->avg_size = safe_division((double)(out->nontrivial_size_sum + trivial), (double)trees);
  out->nontrivial_avg_size = safe_division((double)out->nontrivial_size_sum, (double)out->num_nontrivial_trees);
  out->nontrivial_avg_rank = safe_division((double)out->nontrivial_rank_sum, (double)out->nontrivial_size_sum);
  free(size); free(rsum); free(rmax);
  return level == 4 ? k : 0;
}

/* ------------------------------------------------------------------------ */
/* Graph inputs (bench/input/generation.cpp:60-96, "bfs" processing).        */
/* ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t x);

/* Seeded random undirected edges (not reference code: KaGen is unavailable;
 * the same arithmetic runs on the device, kascade_amd/csrc/graph.hip).
 * kind 0 GNM: endpoints uniform; kind 1 RMAT: one quadrant per bit of
 * ceil(log2 n) bits with 16-bit thresholds for (0.57, 0.19, 0.19, 0.05), ids
 * folded by modulo.  A self loop's head moves to the next vertex. */
void orc_gen_graph(int kind, uint64_t n, uint64_t m, uint64_t seed, uint64_t *src, uint64_t *dst) {
  uint32_t bits = 0;
  while (bits < 63 && (UINT64_C(1) << bits) < n) bits++;
  for (uint64_t e = 0; e < m; e++) {
    uint64_t a, b;
    if (kind == 0) {
      a = orc_splitmix64(seed ^ (2 * e + 1) * UINT64_C(0x9E3779B97F4A7C15)) % n;
      b = orc_splitmix64(seed ^ (2 * e + 2) * UINT64_C(0xD1B54A32D192ED03)) % n;
    } else {
      a = 0;
      b = 0;
      for (uint32_t k = 0; k < bits; k++) {
        uint32_t r = (uint32_t)(orc_splitmix64(seed ^ (e * 64 + k + 1) * UINT64_C(0x94D049BB133111EB)) & 0xFFFF);
        uint32_t q = r < 37356 ? 0 : r < 49807 ? 1 : r < 62259 ? 2 : 3;
        a = (a << 1) | (q >> 1);
        b = (b << 1) | (q & 1);
      }
      a %= n;
      b %= n;
    }
    if (a == b) b = (b + 1) % n;
    src[e] = a;
    dst[e] = b;
  }
}

/* Random geometric graph in the unit square (the role of KaGen's rgg2d
 * generator in the reference's submission suites,
 * experiments/submission/sparse-ruling-set-eulertour.suite.yaml:19-28; not
 * reference code: KaGen is unavailable, and the device generator in
 * kascade_amd/csrc/graph.hip does the same arithmetic).  Point i has 31-bit
 * fixed-point coordinates from the seed; an undirected edge joins i < j when
 * (xi-xj)^2 + (yi-yj)^2 <= r^2 in those units, with r the radius r_fp the
 * caller derives from the expected edge count (orc_rgg2d_radius).  Points are
 * bucketed on a g x g grid of cells at least r wide; edges are emitted grouped
 * by i ascending, then by neighbour cell (rows, then columns, -1..1), then by
 * j ascending.  Writes at most cap edges and returns their number. */
uint64_t orc_rgg2d_radius(uint64_t n, uint64_t m) {
  if (n < 2 || m == 0) return 0;
  double r = sqrt(2.0 * (double)m / (3.141592653589793 * (double)n * (double)n));
  if (r >= 1.0) return UINT64_C(1) << 31;
  return (uint64_t)(r * 2147483648.0);
}

uint64_t orc_rgg2d_cells(uint64_t n, uint64_t r_fp) {
  uint64_t g = (UINT64_C(1) << 31) / r_fp;
  uint64_t s = 1;
  while (s * s < n) s++;
  if (g > s) g = s;
  if (g > 65535) g = 65535;
  return g ? g : 1;
}

uint64_t orc_gen_rgg2d(uint64_t n, uint64_t m, uint64_t seed, uint64_t *src, uint64_t *dst, uint64_t cap) {
  const uint64_t r_fp = orc_rgg2d_radius(n, m);
  if (r_fp == 0) return 0;
  const uint64_t g = orc_rgg2d_cells(n, r_fp);
  const uint64_t w = ((UINT64_C(1) << 31) + g - 1) / g;
  const uint64_t r2 = r_fp * r_fp;
  uint32_t *x = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *y = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t *cell = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint64_t *start = (uint64_t *)calloc(g * g + 1, sizeof(uint64_t));
  uint64_t *order = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; i++) {
    x[i] = (uint32_t)(orc_splitmix64(seed ^ (2 * i + 1) * UINT64_C(0xA0761D6478BD642F)) >> 33);
    y[i] = (uint32_t)(orc_splitmix64(seed ^ (2 * i + 2) * UINT64_C(0xE7037ED1A0B428DB)) >> 33);
    cell[i] = (y[i] / w) * g + x[i] / w;
    start[cell[i] + 1]++;
  }
  for (uint64_t c = 0; c < g * g; c++) start[c + 1] += start[c];
  uint64_t *fill = (uint64_t *)malloc((g * g + 1) * sizeof(uint64_t));
  memcpy(fill, start, (g * g + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; i++) order[fill[cell[i]]++] = i; /* stable: ascending i per cell */
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; i++) {
    const int64_t cx = (int64_t)(x[i] / w), cy = (int64_t)(y[i] / w);
    for (int64_t dy = -1; dy <= 1; dy++) {
      for (int64_t dx = -1; dx <= 1; dx++) {
        const int64_t ax = cx + dx, ay = cy + dy;
        if (ax < 0 || ay < 0 || ax >= (int64_t)g || ay >= (int64_t)g) continue;
        const uint64_t c = (uint64_t)ay * g + (uint64_t)ax;
        for (uint64_t p = start[c]; p < start[c + 1]; p++) {
          const uint64_t j = order[p];
          if (j <= i) continue;
          const uint64_t ddx = x[i] > x[j] ? x[i] - x[j] : x[j] - x[i];
          const uint64_t ddy = y[i] > y[j] ? y[i] - y[j] : y[j] - y[i];
          if (ddx * ddx + ddy * ddy > r2) continue;
          if (k < cap) {
            src[k] = i;
            dst[k] = j;
          }
          k++;
        }
      }
    }
  }
  free(x);
  free(y);
  free(cell);
  free(start);
  free(fill);
  free(order);
  return k;
}

/* generate_bfs_tree, bench/input/bfs_forest_generation.cpp:52-76, on one rank:
 *   handle_isolated_vertices (:40-50): a vertex without neighbours is its own
 *   parent and visited;
 *   then, while an unvisited vertex is left, the smallest one starts a
 *   distributed_bfs (bench/input/bfs.hpp:30-74): the frontier holds (parent,
 *   u) pairs, a pair whose u is unvisited marks u visited with that parent and
 *   appends (u, v) for every neighbour v in adjacency order.
 * The start vertex's parent is itself.  xadj has n + 1 entries. */
void orc_bfs_forest(const uint64_t *xadj, const uint64_t *adj, uint64_t n, uint64_t *parent) {
  uint8_t *visited = (uint8_t *)calloc(n ? n : 1, 1);
  for (uint64_t v = 0; v < n; v++) {
    if (xadj[v] == xadj[v + 1]) {
      visited[v] = 1;
      parent[v] = v;
    }
  }
  uint64_t total = xadj[n];
  uint64_t cap = total + n + 1;
  uint64_t *fp = (uint64_t *)malloc(cap * sizeof(uint64_t)), *fu = (uint64_t *)malloc(cap * sizeof(uint64_t));
  uint64_t *np = (uint64_t *)malloc(cap * sizeof(uint64_t)), *nu = (uint64_t *)malloc(cap * sizeof(uint64_t));
  uint64_t next_start = 0;
  for (;;) {
    while (next_start < n && visited[next_start]) next_start++;
    if (next_start >= n) break;
    uint64_t nf = 1;
    fp[0] = next_start;
    fu[0] = next_start;
    while (nf) {
      uint64_t nn = 0;
      for (uint64_t k = 0; k < nf; k++) {
        uint64_t u = fu[k];
        if (visited[u]) continue;
        visited[u] = 1;
        parent[u] = fp[k];
        for (uint64_t j = xadj[u]; j < xadj[u + 1]; j++) {
          np[nn] = u;
          nu[nn] = adj[j];
          nn++;
        }
      }
      uint64_t *t;
      t = fp; fp = np; np = t;
      t = fu; fu = nu; nu = t;
      nf = nn;
    }
  }
  free(visited); free(fp); free(fu); free(np); free(nu);
}
