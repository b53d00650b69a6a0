/*
 * odp_oracle.c -- CPU oracle for the ASPL/diameter evaluator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (odp_amd/, include/)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may, and only as the checker.
 *
 * It restates, in plain C with 64-bit indices, the method the reference
 * library uses behind ODP_Set_aspl:
 *   - MATRIX         : aspl.c:35-75   (one bit matrix, all sources at once)
 *   - MATRIX_SAVING  : aspl.c:77-119  (column blocks of `chunk` words)
 *   - BFS            : aspl.c:166-221 + utils.c:2060-2100 (serial top-down)
 *   - per-level OR   : utils.c:1815-1861 (matmul) / 1938-1984 (matmul_CHUNK)
 *   - symmetry index : utils.c:67-88 (ODP_ROTATE), 195-210 (NORM /
 *                      GLOBAL_ADJ_GENERAL), 213-234 (ODP_LOCAL_INDEX_GRID),
 *                      237-274 (GLOBAL_ADJ_GRID), 2103-2124 (ODP_Create_itable)
 *   - Init variants  : aspl.c:223-288 (num_degrees expansion per variant)
 *   - sentinels      : aspl.c:332-336
 *   - one rank's share of the MPI+CUDA flavour: aspl_mpi_cuda.cu:44-123
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   - all three kinds on s=1 graphs: the reference's golden outputs
 *     sample/graph/{general,grid}/ASPL.txt, closed forms (K_n, C_n, Petersen,
 *     Q_k, path), networkx all-pairs shortest paths on random graphs, and
 *     the reference library itself (oracle/_ref) on generated graphs;
 *   - symmetric general / grid paths (s = 2, 4): the invariant "result at s
 *     equals result at s=1 on the expanded graph", plus the values the
 *     reference produced for the symmetric fixtures n24d4g4, w6h6d3r2g2,
 *     w6h6d3r2g4, w10h10d4r4g4 (tests/golden/symmetric_fixtures.json), plus
 *     oracle/_ref on generated symmetric graphs;
 *   - MATRIX vs MATRIX_SAVING vs BFS: bit-identical for every chunk width;
 *   - rank shares: the ranks' partials combine to the whole evaluation, and
 *     each rank's partial equals breadth-first searches from its own sources
 *     (sum of distances minus n-1 per source, times s; the largest eccentricity).
 * oracle_set_fault() plants known mistakes so the tests can show each check
 * above rejects them.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <float.h>

#define ORACLE_MATRIX        1
#define ORACLE_MATRIX_SAVING 2
#define ORACLE_BFS           3

/* Planted mistakes, used only by the tests that prove the checks have teeth. */
enum {
  FAULT_NONE = 0,
  FAULT_SAVING_DROP_S = 1,    /* forget the x symmetries factor in MATRIX_SAVING */
  FAULT_DIAMETER_OFF = 2,     /* diameter one too small                          */
  FAULT_ROTATE_SWAP = 3,      /* swap the 90 and 270 degree rotations             */
  FAULT_CIRCULANT_SHIFT = 4,  /* wrong orbit shift in the symmetric general case  */
  FAULT_SKIP_LAST_NBR = 5,    /* drop the last neighbour of every row             */
  FAULT_ASPL_DENOM = 6,       /* divide by n*n instead of n*(n-1)                 */
  FAULT_SHARD_CHUNK = 7       /* MATRIX shards: floor(W / procs) words, not ceil   */
};
static int g_fault = FAULT_NONE;
void oracle_set_fault(int code) { g_fault = code; }

/* ---- symmetry index math (utils.c:67-88, 195-274, 2103-2124) ---------- */

static int o_rotate(int v, int width, int height, int symmetries, int angle)
{
  int w = v / height, h = v % height;
  if (symmetries == 2) return (width - w - 1) * height + (height - h - 1);
  if (g_fault == FAULT_ROTATE_SWAP) angle = (angle == 90) ? 270 : (angle == 270 ? 90 : angle);
  if (angle == 90)  return h * height + (height - w - 1);
  if (angle == 180) return (height - w - 1) * height + (height - h - 1);
  return (height - h - 1) * height + w; /* 270 */
}

static int o_local_index_grid(int x, int height, int symmetries)
{
  if (symmetries != 4) return x;
  int bh = height / 2;
  return (x / height) * bh + (x % height);
}

/* neighbour d of vertex v in a 2- or 4-fold symmetric grid whose table holds
 * only the first-quadrant (s=4) or left-half (s=2) vertices */
static int o_global_adj_grid(int width, int height, int degree, int symmetries,
                             const int *adj, int v, int d)
{
  int bw = width / 2, bh = height / 2, w = v / height, h = v % height;
  if (symmetries == 1) return adj[(int64_t)v * degree + d];
  if (symmetries == 2) {
    if (w < bw) return adj[(int64_t)v * degree + d];
    int y = adj[(int64_t)o_rotate(v, width, height, 2, 180) * degree + d];
    return o_rotate(y, width, height, 2, 180);
  }
  if (w < bw && h < bh)
    return adj[(int64_t)o_local_index_grid(v, height, 4) * degree + d];
  int to_base, back;
  if (w < bw)       { to_base = 270; back = 90;  }  /* upper-left quadrant  */
  else if (h >= bh) { to_base = 180; back = 180; }  /* upper-right quadrant */
  else              { to_base = 90;  back = 270; }  /* lower-right quadrant */
  int x = o_rotate(v, width, height, 4, to_base);
  int li = o_local_index_grid(x, height, 4), based = width * height / 4;
  if (li < 0 || li >= based) li = ((li % based) + based) % based;   /* only under a planted fault */
  int y = adj[(int64_t)li * degree + d];
  int r = o_rotate(y, width, height, 4, back), n = width * height;
  return (r < 0 || r >= n) ? ((r % n) + n) % n : r;
}

static void o_create_itable(int width, int height, int symmetries, int *itable)
{
  int based = (width * height) / symmetries;
  if (symmetries == 2) {
    for (int i = 0; i < based; i++) {
      itable[i] = i;
      itable[o_rotate(i, width, height, 2, 180)] = i + based;
    }
  } else {
    int bh = height / 2;
    for (int i = 0; i < based; i++) {
      int v = (i / bh) * height + (i % bh);
      itable[v] = i;
      itable[o_rotate(v, width, height, 4, 90)]  = i + based;
      itable[o_rotate(v, width, height, 4, 180)] = i + 2 * based;
      itable[o_rotate(v, width, height, 4, 270)] = i + 3 * based;
    }
  }
}

/* ---- evaluator state (mirrors the statics of aspl.c:2-9) --------------- */

typedef struct {
  int nodes, degree, symmetries, height, grid_s;
  int *num_degrees;   /* expanded to length nodes, or NULL (regular) */
  int *itable;        /* grid s in {2,4} only                          */
} octx;

/* neighbour j of matrix row i, returned as a row index and the row's degree.
 * general s>1: utils.c:1849-1854; grid s>1: utils.c:1835-1839 */
static int o_deg(const octx *c, int i)
{
  int d = c->num_degrees ? c->num_degrees[i] : c->degree;
  if (g_fault == FAULT_SKIP_LAST_NBR && d > 0) d--;
  return d;
}

static int o_nbr_general(const octx *c, const int *adj, int i, int j)
{
  int based = c->nodes / c->symmetries;
  int p = i / based, m = i - p * based;
  int shift = p * based;
  if (g_fault == FAULT_CIRCULANT_SHIFT && p > 0) shift = (p - 1) * based;
  int n = adj[(int64_t)m * c->degree + j] + shift;
  if (n >= c->nodes) n -= c->nodes;
  return n;
}

/* One BFS level of the bit matrix:  B[row(i)] |= OR_j A[row(nbr(i,j))]
 * over `words` 64-bit words per row (utils.c:1815-1861, 1938-1984). */
static void o_level(const octx *c, const int *adj, const uint64_t *A, uint64_t *B, int64_t words)
{
  int n = c->nodes;
  if (c->grid_s) {
    int width = n / c->height;
    for (int i = 0; i < n; i++) {
      int64_t ii = c->itable[i];
      int d = o_deg(c, i);
      for (int j = 0; j < d; j++) {
        int v = o_global_adj_grid(width, c->height, c->degree, c->symmetries, adj, i, j);
        int64_t nn = c->itable[v];
        for (int64_t k = 0; k < words; k++) B[ii * words + k] |= A[nn * words + k];
      }
    }
  } else {
    for (int i = 0; i < n; i++) {
      int d = o_deg(c, i);
      for (int j = 0; j < d; j++) {
        int64_t v = (c->symmetries == 1) ? adj[(int64_t)i * c->degree + j] : o_nbr_general(c, adj, i, j);
        for (int64_t k = 0; k < words; k++) B[(int64_t)i * words + k] |= A[v * words + k];
      }
    }
  }
}

static uint64_t o_popcount(const uint64_t *B, int64_t len)
{
  uint64_t num = 0;
  for (int64_t i = 0; i < len; i++) num += (uint64_t)__builtin_popcountll(B[i]);
  return num;
}

/* MATRIX, aspl.c:35-75 */
static void o_matrix(const octx *c, const int *adj, int *diameter, long *sum, double *ASPL)
{
  int n = c->nodes, s = c->symmetries;
  int64_t W = ((int64_t)(n / s) + 63) / 64;
  uint64_t *A = calloc((size_t)(n * W), 8), *B = calloc((size_t)(n * W), 8);
  for (int i = 0; i < n / s; i++) {
    A[i * W + i / 64] = B[i * W + i / 64] = 1ULL << (i % 64);
  }
  *sum = (long)n * (n - 1);
  *diameter = 1;
  for (int kk = 0; kk < n; kk++) {
    o_level(c, adj, A, B, W);
    uint64_t num = o_popcount(B, (int64_t)n * W) * (uint64_t)s;
    if (num == (uint64_t)n * n) break;
    uint64_t *t = A; A = B; B = t;
    *sum += (long)n * n - (long)num;
    *diameter += 1;
  }
  free(A); free(B);
  double denom = (g_fault == FAULT_ASPL_DENOM) ? (double)n * n : ((double)n - 1) * n;
  *ASPL = *sum / denom;
  *sum /= 2.0;
}

/* MATRIX_SAVING, aspl.c:77-119, with an arbitrary chunk width in words */
static void o_matrix_saving(const octx *c, const int *adj, int chunk, int *diameter, long *sum, double *ASPL)
{
  int n = c->nodes, s = c->symmetries, based = n / s;
  int64_t W = ((int64_t)based + 63) / 64;
  int64_t parsize = (W + chunk - 1) / chunk;
  uint64_t *A = malloc((size_t)n * chunk * 8), *B = malloc((size_t)n * chunk * 8);
  *sum = (long)n * (n - 1);
  *diameter = 1;
  for (int64_t t = 0; t < parsize; t++) {
    memset(A, 0, (size_t)n * chunk * 8);
    memset(B, 0, (size_t)n * chunk * 8);
    int64_t l;
    for (l = 0; l < 64LL * chunk && 64LL * t * chunk + l < based; l++) {
      int64_t off = (64LL * t * chunk + l) * chunk + l / 64;
      A[off] = B[off] = 1ULL << (l % 64);
    }
    int kk;
    for (kk = 0; kk < n; kk++) {
      o_level(c, adj, A, B, chunk);
      uint64_t num = o_popcount(B, (int64_t)n * chunk);
      if (num == (uint64_t)n * (uint64_t)l) break;
      uint64_t *tmp = A; A = B; B = tmp;
      long mult = (g_fault == FAULT_SAVING_DROP_S) ? 1 : s;
      *sum += ((long)n * l - (long)num) * mult;
    }
    if (kk + 1 > *diameter) *diameter = kk + 1;
  }
  free(A); free(B);
  double denom = (g_fault == FAULT_ASPL_DENOM) ? (double)n * n : ((double)n - 1) * n;
  *ASPL = *sum / denom;
  *sum /= 2.0;
}

/* BFS, aspl.c:166-221 with the serial top-down step of utils.c:2060-2100 */
static void o_bfs(const octx *c, const int *adj, int *diameter, long *sum, double *ASPL)
{
  int n = c->nodes, s = c->symmetries, based = n / s;
  int *frontier = malloc(sizeof(int) * n), *next = malloc(sizeof(int) * n);
  int *dist = malloc(sizeof(int) * n);
  char *seen = malloc((size_t)n);
  *diameter = 0;
  *sum = 0;
  int width = c->grid_s ? n / c->height : 0;
  for (int src = 0; src < based; src++) {
    memset(seen, 0, (size_t)n);
    int root = src;
    if (c->grid_s && s == 4) {
      int bh = c->height / 2;
      root = (src / bh) * c->height + (src % bh);
    }
    frontier[0] = root; dist[root] = 0; seen[root] = 1;
    int nf = 1, level = 0;
    while (1) {
      int cnt = 0;
      for (int f = 0; f < nf; f++) {
        int v = frontier[f];
        if (c->grid_s) {
          int d = o_deg(c, v);
          for (int j = 0; j < d; j++) {
            int u = o_global_adj_grid(width, c->height, c->degree, s, adj, v, j);
            if (!seen[u]) { seen[u] = 1; dist[u] = level; next[cnt++] = u; }
          }
        } else {
          int p = v / based, m = v - p * based;
          int d = o_deg(c, m);
          for (int j = 0; j < d; j++) {
            int u = adj[(int64_t)m * c->degree + j] + p * based;
            if (u >= n) u -= n;
            if (!seen[u]) { seen[u] = 1; dist[u] = level; next[cnt++] = u; }
          }
        }
      }
      level++;
      if (cnt == 0) break;
      int *t = frontier; frontier = next; next = t; nf = cnt;
    }
    if (level - 1 > *diameter) *diameter = level - 1;
    if (src == 0) {
      for (int i = 1; i < n; i++)
        if (!seen[i]) { *diameter = INT_MAX; goto out; }
    }
    for (int i = 0; i < n; i++) *sum += (long)(dist[i] + 1) * s;
  }
  *sum = (*sum - n) / 2;
  *ASPL = *sum / (((double)n - 1) * n) * 2;
out:
  free(frontier); free(next); free(dist); free(seen);
}

/*
 * One full Init / Set / Finalize cycle.
 *   family 0 = general (ODP_Init_aspl_general[_s]),  1 = grid (ODP_Init_aspl_grid[_s])
 *   nodes_or_width: nodes (general) or width (grid); height: grid height (ignored for general)
 *   num_degrees: NULL, or length n (s == 1) / n/s (s > 1) as the _s entry points take it
 *   adjacency: int32[n/s][degree]
 *   kind: 1 MATRIX, 2 MATRIX_SAVING, 3 BFS;  chunk: words per block for MATRIX_SAVING
 * Returns 0, or -1 on invalid parameters (where the reference exits).
 */
int oracle_eval(int family, int nodes_or_width, int height, int degree, const int *num_degrees,
                int symmetries, const int *adjacency, int kind, int chunk,
                int *diameter, long *sum, double *ASPL)
{
  octx c;
  memset(&c, 0, sizeof c);
  int width = nodes_or_width;
  int nodes = family ? width * height : nodes_or_width;
  if (symmetries < 1 || nodes % symmetries != 0 || chunk < 1) return -1;
  c.nodes = nodes; c.degree = degree; c.symmetries = symmetries;
  c.height = family ? height : -1;
  int based = nodes / symmetries;
  if (family && symmetries > 1) {
    if (symmetries != 2 && symmetries != 4) return -1;
    if (symmetries == 2 && width % 2) return -1;
    if (symmetries == 4 && (width % 2 || height % 2 || width != height)) return -1;
    c.grid_s = 1;
    c.itable = calloc((size_t)nodes, sizeof(int));  /* zeroed: a planted rotation fault may leave holes */
    o_create_itable(width, height, symmetries, c.itable);
  }
  if (num_degrees) {
    c.num_degrees = malloc(sizeof(int) * nodes);
    if (symmetries == 1) {
      memcpy(c.num_degrees, num_degrees, sizeof(int) * nodes);
    } else if (!family) {            /* aspl.c:228-244 */
      for (int i = 0; i < symmetries; i++)
        for (int j = 0; j < based; j++) c.num_degrees[i * based + j] = num_degrees[j];
    } else if (symmetries == 2) {    /* aspl.c:267-271 */
      for (int i = 0; i < based; i++) {
        c.num_degrees[i] = num_degrees[i];
        c.num_degrees[o_rotate(i, width, height, 2, 180)] = num_degrees[i];
      }
    } else {                         /* aspl.c:273-281 */
      for (int i = 0; i < based; i++) {
        int v = (i / (height / 2)) * height + (i % (height / 2));
        c.num_degrees[v] = num_degrees[i];
        c.num_degrees[o_rotate(v, width, height, 4, 90)]  = num_degrees[i];
        c.num_degrees[o_rotate(v, width, height, 4, 180)] = num_degrees[i];
        c.num_degrees[o_rotate(v, width, height, 4, 270)] = num_degrees[i];
      }
    }
  }
  if (kind == ORACLE_MATRIX)             o_matrix(&c, adjacency, diameter, sum, ASPL);
  else if (kind == ORACLE_MATRIX_SAVING) o_matrix_saving(&c, adjacency, chunk, diameter, sum, ASPL);
  else                                   o_bfs(&c, adjacency, diameter, sum, ASPL);
  if (g_fault == FAULT_DIAMETER_OFF && *diameter <= nodes) *diameter -= 1;
  if (*diameter > nodes) {                /* aspl.c:332-336 */
    *diameter = INT_MAX;
    *sum = LONG_MAX;
    *ASPL = DBL_MAX;
  }
  free(c.num_degrees);
  free(c.itable);
  return 0;
}

/* The per-rank loop of the MPI+CUDA flavour (aspl_mpi_cuda.cu:44-80 and
 * 90-123): column blocks t = rank, rank + procs, ... of `chunk` words each,
 * every block's levels until it reaches all n x l pairs (or n levels), the
 * rank-local sum += (n l - num) s starting at 0 and diameter = max(kk + 1)
 * starting at 1 -- what each rank holds before the two Allreduces at :82-83. */
static void o_rank_blocks(const octx *c, const int *adj, int64_t chunk, int rank, int procs, long *sum, int *diameter)
{
  int n = c->nodes, s = c->symmetries, based = n / s;
  int64_t W = ((int64_t)based + 63) / 64;
  int64_t parsize = (W + chunk - 1) / chunk;
  uint64_t *A = malloc((size_t)n * chunk * 8), *B = malloc((size_t)n * chunk * 8);
  *sum = 0;
  *diameter = 1;
  for (int64_t t = rank; t < parsize; t += procs) {
    memset(A, 0, (size_t)n * chunk * 8);
    memset(B, 0, (size_t)n * chunk * 8);
    int64_t l;
    for (l = 0; l < 64LL * chunk && 64LL * t * chunk + l < based; l++) {
      int64_t off = (64LL * t * chunk + l) * chunk + l / 64;
      A[off] = B[off] = 1ULL << (l % 64);
    }
    int kk;
    for (kk = 0; kk < n; kk++) {
      o_level(c, adj, A, B, chunk);
      uint64_t num = o_popcount(B, (int64_t)n * chunk);
      if (num == (uint64_t)n * (uint64_t)l) break;
      uint64_t *tmp = A; A = B; B = tmp;
      *sum += ((long)n * l - (long)num) * s;
    }
    if (kk + 1 > *diameter) *diameter = kk + 1;
  }
  free(A); free(B);
}

/*
 * One rank's share of ODP_Set_aspl in the MPI+CUDA flavour, before the
 * combine: MATRIX takes ceil(W / procs)-word blocks, rank r the block t = r
 * (aspl_mpi_cuda.cu:47-54, W = ceil((n/s)/64)); MATRIX_SAVING takes the
 * 64-word blocks t = r, r + procs, ... (GPU_CHUNK, :92-94).  sum is the
 * rank-local excess over the n(n-1) base (the reference's *sum before the
 * Allreduce, :83), diameter the rank-local maximum of kk + 1 (n + 1 when a
 * block never completes).  The combine is :82-87: diameter = max over ranks,
 * sum = sum over ranks + n(n-1), ASPL = sum / ((n-1) n), sum /= 2.
 * Symmetric grids use the reference's row order (utils.c:2103-2124).
 * Returns 0, or -1 on invalid parameters.
 */
int oracle_eval_shard(int family, int nodes_or_width, int height, int degree, const int *num_degrees,
                      int symmetries, const int *adjacency, int kind, int rank, int procs,
                      long *sum, int *diameter)
{
  octx c;
  memset(&c, 0, sizeof c);
  int width = nodes_or_width;
  int nodes = family ? width * height : nodes_or_width;
  if (symmetries < 1 || nodes % symmetries != 0 || procs < 1 || rank < 0 || rank >= procs) return -1;
  if (kind != ORACLE_MATRIX && kind != ORACLE_MATRIX_SAVING) return -1;
  c.nodes = nodes; c.degree = degree; c.symmetries = symmetries;
  c.height = family ? height : -1;
  int based = nodes / symmetries;
  if (family && symmetries > 1) {
    if (symmetries != 2 && symmetries != 4) return -1;
    c.grid_s = 1;
    c.itable = calloc((size_t)nodes, sizeof(int));
    o_create_itable(width, height, symmetries, c.itable);
  }
  if (num_degrees) {
    c.num_degrees = malloc(sizeof(int) * nodes);
    if (symmetries == 1) {
      memcpy(c.num_degrees, num_degrees, sizeof(int) * nodes);
    } else if (!family) {
      for (int i = 0; i < symmetries; i++)
        for (int j = 0; j < based; j++) c.num_degrees[i * based + j] = num_degrees[j];
    } else if (symmetries == 2) {
      for (int i = 0; i < based; i++) {
        c.num_degrees[i] = num_degrees[i];
        c.num_degrees[o_rotate(i, width, height, 2, 180)] = num_degrees[i];
      }
    } else {
      for (int i = 0; i < based; i++) {
        int v = (i / (height / 2)) * height + (i % (height / 2));
        c.num_degrees[v] = num_degrees[i];
        c.num_degrees[o_rotate(v, width, height, 4, 90)]  = num_degrees[i];
        c.num_degrees[o_rotate(v, width, height, 4, 180)] = num_degrees[i];
        c.num_degrees[o_rotate(v, width, height, 4, 270)] = num_degrees[i];
      }
    }
  }
  int64_t W = ((int64_t)based + 63) / 64;
  int64_t chunk = kind == ORACLE_MATRIX ? (W + procs - 1) / procs : 64;
  if (g_fault == FAULT_SHARD_CHUNK && kind == ORACLE_MATRIX) chunk = W / procs > 0 ? W / procs : 1;
  o_rank_blocks(&c, adjacency, chunk, rank, procs, sum, diameter);
  free(c.num_degrees);
  free(c.itable);
  return 0;
}

/* Expands a symmetric table to the full s=1 adjacency [nodes][degree] with
 * global vertex ids (general: utils.c:991-1010; grid: utils.c:152-193).
 * num_degrees (length n/s or NULL) gives the valid entries per base row;
 * entries beyond it are written as -1. */
void oracle_expand(int family, int nodes_or_width, int height, int degree, const int *num_degrees,
                   int symmetries, const int *adjacency, int *out)
{
  int width = nodes_or_width;
  int nodes = family ? width * height : nodes_or_width;
  int based = nodes / symmetries;
  for (int64_t i = 0; i < (int64_t)nodes * degree; i++) out[i] = -1;
  if (!family || symmetries == 1) {
    for (int p = 0; p < symmetries; p++)
      for (int j = 0; j < based; j++) {
        int d = num_degrees ? num_degrees[j] : degree;
        for (int k = 0; k < d; k++) {
          int v = adjacency[(int64_t)j * degree + k] + p * based;
          while (v >= nodes) v -= nodes;
          out[(int64_t)(p * based + j) * degree + k] = v;
        }
      }
    return;
  }
  for (int v = 0; v < nodes; v++) {
    int bw = width / 2, bh = height / 2, w = v / height, h = v % height, base;
    if (symmetries == 2) base = (w < bw) ? v : o_rotate(v, width, height, 2, 180);
    else if (w < bw && h < bh) base = o_local_index_grid(v, height, 4);
    else if (w < bw)  base = o_local_index_grid(o_rotate(v, width, height, 4, 270), height, 4);
    else if (h >= bh) base = o_local_index_grid(o_rotate(v, width, height, 4, 180), height, 4);
    else              base = o_local_index_grid(o_rotate(v, width, height, 4, 90), height, 4);
    int d = num_degrees ? num_degrees[base] : degree;
    for (int k = 0; k < d; k++)
      out[(int64_t)v * degree + k] = o_global_adj_grid(width, height, degree, symmetries, adjacency, v, k);
  }
}
