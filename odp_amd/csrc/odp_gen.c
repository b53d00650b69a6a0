/*
 * odp_gen.c -- seeded synthetic graphs for the ASPL evaluator, plus the 2-opt
 * edge swap an annealing chain applies between evaluations.
 *
 * The reference generates graphs with glibc rand() (utils.c:1036-1114,
 * 1577-1749); these generators are this library's own, deterministic across
 * hosts (splitmix64-seeded xoshiro256**), and always return SIMPLE graphs
 * (no self-loops, no multi-edges), which the reference does not guarantee
 * (its multi-edge check is disabled, utils.c:469-498).
 *
 * Tables produced here use the reference's adjacency layout: int32
 * [rows][degree], row-major, global vertex ids (README.md:122-150).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "odpx.h"

/* ---- PRNG: splitmix64 -> xoshiro256** ----------------------------------- */

static uint64_t splitmix64(uint64_t *x)
{
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void ODPX_Rng_seed(uint64_t seed, uint64_t state[4])
{
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) state[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t ODPX_Rng_next(uint64_t s[4])
{
  uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
  return r;
}

/* uniform in [0, bound), bound >= 1 (Lemire's multiply-shift, rejection-free
 * bias < 2^-32 for the bounds used here) */
static inline uint32_t rnd(uint64_t s[4], uint32_t bound)
{
  return (uint32_t)(((ODPX_Rng_next(s) >> 32) * (uint64_t)bound) >> 32);
}

/* ---- edge multiset keyed by the unordered pair -------------------------- */

typedef struct { uint64_t *key; int32_t *cnt; uint64_t mask; } eset;

static int eset_init(eset *h, int64_t nedges)
{
  uint64_t cap = 16;
  while (cap < (uint64_t)nedges * 2 + 16) cap <<= 1;
  h->key = malloc(cap * sizeof(uint64_t));
  h->cnt = calloc(cap, sizeof(int32_t));
  h->mask = cap - 1;
  if (!h->key || !h->cnt) return -1;
  memset(h->key, 0xff, cap * sizeof(uint64_t));
  return 0;
}
static void eset_free(eset *h) { free(h->key); free(h->cnt); }
static inline uint64_t ekey(int a, int b)
{
  uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  return ((uint64_t)lo << 32) | hi;
}
static inline uint64_t ehome(const eset *h, uint64_t k) { return ((k * 0x9E3779B97F4A7C15ULL) >> 20) & h->mask; }
static inline uint64_t eslot(const eset *h, uint64_t k)
{
  uint64_t i = ehome(h, k);
  while (h->key[i] != k && h->key[i] != ~0ULL) i = (i + 1) & h->mask;
  return i;
}
/* linear-probing delete with backward shift, so the table never fills up with
 * dead keys however many edges come and go */
static void edel(eset *h, uint64_t i)
{
  uint64_t j = i;
  for (;;) {
    j = (j + 1) & h->mask;
    if (h->key[j] == ~0ULL) break;
    uint64_t k = ehome(h, h->key[j]);
    int stays = (i <= j) ? (i < k && k <= j) : (i < k || k <= j);
    if (!stays) { h->key[i] = h->key[j]; h->cnt[i] = h->cnt[j]; i = j; }
  }
  h->key[i] = ~0ULL;
  h->cnt[i] = 0;
}
static inline int eget(const eset *h, int a, int b) { uint64_t i = eslot(h, ekey(a, b)); return h->key[i] == ~0ULL ? 0 : h->cnt[i]; }
static inline void eadd(eset *h, int a, int b, int d)
{
  uint64_t k = ekey(a, b), i = eslot(h, k);
  if (h->key[i] == ~0ULL) h->key[i] = k;
  h->cnt[i] += d;
  if (h->cnt[i] <= 0) edel(h, i);
}

/* ---- connectivity -------------------------------------------------------- */

/* labels[] = component id per vertex; returns the number of components */
static int components(int n, int degree, const int *adj, int *label, int *queue)
{
  for (int i = 0; i < n; i++) label[i] = -1;
  int nc = 0;
  for (int s = 0; s < n; s++) {
    if (label[s] >= 0) continue;
    int head = 0, tail = 0;
    queue[tail++] = s; label[s] = nc;
    while (head < tail) {
      int v = queue[head++];
      for (int j = 0; j < degree; j++) {
        int u = adj[(int64_t)v * degree + j];
        if (u >= 0 && label[u] < 0) { label[u] = nc; queue[tail++] = u; }
      }
    }
    nc++;
  }
  return nc;
}

/* ---- random d-regular simple connected graph ---------------------------- */

/* Pairing model: nodes*degree stubs shuffled and paired; self-loops and
 * repeated pairs are then repaired by swapping endpoints with random other
 * pairs, and components are joined the same way. */
int ODPX_Generate_regular(int nodes, int degree, uint64_t seed, int *adjacency)
{
  if (nodes < 2 || degree < 1 || degree >= nodes || ((int64_t)nodes * degree) % 2) return -1;
  int64_t nst = (int64_t)nodes * degree, ne = nst / 2;
  uint64_t rs[4];
  ODPX_Rng_seed(seed, rs);
  int *st = malloc(sizeof(int) * nst);
  if (!st) return -1;
  for (int64_t i = 0; i < nst; i++) st[i] = (int)(i / degree);
  for (int64_t i = nst - 1; i > 0; i--) {
    int64_t j = (int64_t)(((ODPX_Rng_next(rs) >> 11) % (uint64_t)(i + 1)));
    int t = st[i]; st[i] = st[j]; st[j] = t;
  }
  /* pair i = (st[2i], st[2i+1]) */
  eset h;
  if (eset_init(&h, ne)) { free(st); return -1; }
  for (int64_t i = 0; i < ne; i++) eadd(&h, st[2 * i], st[2 * i + 1], 1);
  int guard = 0;
  for (;;) {
    int64_t bad = 0;
    for (int64_t i = 0; i < ne; i++) {
      int a = st[2 * i], b = st[2 * i + 1];
      if (a != b && eget(&h, a, b) == 1) continue;
      bad++;
      for (int tries = 0; tries < 1000; tries++) {
        int64_t j = (int64_t)((ODPX_Rng_next(rs) >> 11) % (uint64_t)ne);
        if (j == i) continue;
        int c = st[2 * j], d = st[2 * j + 1];
        if (ODPX_Rng_next(rs) & 1) { int t = c; c = d; d = t; }
        /* new pairs (a,c), (b,d) */
        if (a == c || b == d) continue;
        if (ekey(a, c) == ekey(b, d)) continue;
        if (eget(&h, a, c) || eget(&h, b, d)) continue;
        eadd(&h, a, b, -1); eadd(&h, st[2 * j], st[2 * j + 1], -1);
        st[2 * i + 1] = c; st[2 * j] = b; st[2 * j + 1] = d;
        eadd(&h, a, c, 1); eadd(&h, b, d, 1);
        break;
      }
    }
    if (!bad) break;
    if (++guard > 1000) { eset_free(&h); free(st); return -2; }
  }
  int *deg = calloc((size_t)nodes, sizeof(int));
  for (int64_t i = 0; i < ne; i++) {
    int a = st[2 * i], b = st[2 * i + 1];
    adjacency[(int64_t)a * degree + deg[a]++] = b;
    adjacency[(int64_t)b * degree + deg[b]++] = a;
  }
  /* join components with 2-opt swaps between one edge inside vertex 0's
   * component and one outside it */
  int *label = malloc(sizeof(int) * nodes), *queue = malloc(sizeof(int) * nodes);
  int rc = 0;
  for (int round = 0; components(nodes, degree, adjacency, label, queue) > 1; round++) {
    if (round > 100000) { rc = -3; break; }
    int a = (int)rnd(rs, nodes), c = (int)rnd(rs, nodes);
    if (label[a] == label[c]) continue;
    int ja = (int)rnd(rs, degree), jc = (int)rnd(rs, degree);
    int b = adjacency[(int64_t)a * degree + ja], d = adjacency[(int64_t)c * degree + jc];
    if (eget(&h, a, c) || eget(&h, b, d)) continue;
    ODPX_Swap_edges(nodes, degree, adjacency, a, b, c, d);
    eadd(&h, a, b, -1); eadd(&h, c, d, -1); eadd(&h, a, c, 1); eadd(&h, b, d, 1);
  }
  free(label); free(queue); free(deg);
  eset_free(&h); free(st);
  return rc;
}

/* ---- 2-opt ---------------------------------------------------------------- */

static int replace_in_row(int degree, int *row, int from, int to)
{
  for (int j = 0; j < degree; j++)
    if (row[j] == from) { row[j] = to; return j; }
  return -1;
}

/* Replaces edges a-b and c-d by a-c and b-d in a full [n][degree] table. */
int ODPX_Swap_edges(int nodes, int degree, int *adj, int a, int b, int c, int d)
{
  (void)nodes;
  if (replace_in_row(degree, adj + (int64_t)a * degree, b, c) < 0) return -1;
  if (replace_in_row(degree, adj + (int64_t)b * degree, a, d) < 0) return -1;
  if (replace_in_row(degree, adj + (int64_t)c * degree, d, a) < 0) return -1;
  if (replace_in_row(degree, adj + (int64_t)d * degree, c, b) < 0) return -1;
  return 0;
}

static int has_edge(int degree, const int *adj, int a, int b)
{
  const int *row = adj + (int64_t)a * degree;
  for (int j = 0; j < degree; j++) if (row[j] == b) return 1;
  return 0;
}

/* One random 2-opt move on a simple graph: a-b, c-d -> a-c, b-d, redrawn until
 * it creates neither a self-loop nor a multi-edge (the move the annealing
 * driver of sample/SA/general.c:160 applies through ODP_Mutate_adjacency_*).
 * The four touched vertices are returned in swap[4] = {a, b, c, d}. */
int ODPX_Two_opt(int nodes, int degree, int *adj, uint64_t rng[4], int swap[4])
{
  for (int tries = 0; tries < 1000000; tries++) {
    int a = (int)rnd(rng, nodes), c = (int)rnd(rng, nodes);
    int b = adj[(int64_t)a * degree + rnd(rng, degree)];
    int d = adj[(int64_t)c * degree + rnd(rng, degree)];
    if (b < 0 || d < 0) continue;
    if (a == c || a == d || b == c || b == d) continue;
    if (has_edge(degree, adj, a, c) || has_edge(degree, adj, b, d)) continue;
    if (ODPX_Swap_edges(nodes, degree, adj, a, b, c, d)) return -1;
    if (swap) { swap[0] = a; swap[1] = b; swap[2] = c; swap[3] = d; }
    return 0;
  }
  return -1;
}

/* ---- rotationally symmetric grid ------------------------------------------ */

/* 90/180/270-degree rotations of vertex v = x*height + y (utils.c:67-88) */
static inline int grid_rot(int v, int width, int height, int s, int k)
{
  int w = v / height, h = v % height;
  if (s == 2) return (width - w - 1) * height + (height - h - 1);
  for (int i = 0; i < k; i++) {
    int nv = h * height + (height - w - 1);
    w = nv / height; h = nv % height;
  }
  return w * height + h;
}
static inline int grid_dist(int u, int v, int height)
{
  int du = u / height - v / height, dv = u % height - v % height;
  return (du < 0 ? -du : du) + (dv < 0 ? -dv : dv);
}
static inline int orbit_rep(int v, int width, int height, int s)
{
  int best = v;
  int cnt = (s == 4) ? 4 : s;
  for (int k = 1; k < cnt; k++) {
    int r = grid_rot(v, width, height, s, k);
    if (r < best) best = r;
  }
  return best;
}

/*
 * Random `degree`-regular simple graph on a width x height grid whose edges
 * have Manhattan length <= `length` and which is invariant under rotation by
 * 360/s degrees (s in {1, 2, 4}; s = 4 needs width == height, both even;
 * height must be even).  Start: each vertex is tied to a domino partner by
 * `degree` parallel edges (the orbit of vertical dominoes in the base region).
 * Then symmetric 2-opt moves over whole orbits (a-b, c-d -> a-c, b-d applied
 * to every rotation) are accepted when they keep the length bound and do not
 * increase the number of parallel edges; they run until no parallel edge is
 * left, for at least `iters_per_edge` moves per edge orbit, and until the
 * graph is connected.  Output: the full table int32[width*height][degree].
 */
int ODPX_Generate_grid(int width, int height, int degree, int length, int symmetries,
                       uint64_t seed, int iters_per_edge, int *adj)
{
  int s = symmetries, n = width * height;
  if (s != 1 && s != 2 && s != 4) return -1;
  if (height % 2 || (s == 2 && width % 2) || (s == 4 && (width != height || width % 2))) return -1;
  if (degree < 1 || length < 1) return -1;
  int orb = s;
  uint64_t rs[4];
  ODPX_Rng_seed(seed, rs);
  /* dominoes: vertical pairs (x, 2j)-(x, 2j+1) in the base region, rotated */
  int bw = (s == 1) ? width : width / 2, bh = (s == 4) ? height / 2 : height;
  if (bh % 2) return -1;
  for (int x = 0; x < bw; x++)
    for (int y = 0; y < bh; y += 2) {
      int a = x * height + y, b = a + 1;
      for (int k = 0; k < orb; k++) {
        int ra = (k == 0) ? a : grid_rot(a, width, height, s, k);
        int rb = (k == 0) ? b : grid_rot(b, width, height, s, k);
        for (int j = 0; j < degree; j++) {
          adj[(int64_t)ra * degree + j] = rb;
          adj[(int64_t)rb * degree + j] = ra;
        }
      }
    }
  eset h;
  if (eset_init(&h, (int64_t)n * degree / 2)) return -1;
  int64_t defects = 0;
  for (int v = 0; v < n; v++)
    for (int j = 0; j < degree; j++) {
      int u = adj[(int64_t)v * degree + j];
      if (v < u) { if (eget(&h, v, u)) defects++; eadd(&h, v, u, 1); }
    }
  int64_t orbits = (int64_t)n * degree / 2 / orb;
  int64_t min_moves = orbits * (int64_t)iters_per_edge, moves = 0, proposals = 0;
  int *label = malloc(sizeof(int) * n), *queue = malloc(sizeof(int) * n);
  int L = length, rc = 0;
  for (;;) {
    if (defects == 0 && moves >= min_moves) {
      if (components(n, degree, adj, label, queue) == 1) break;
      min_moves += orbits;   /* keep mixing until connected */
    }
    if (++proposals > 4000LL * orbits + 100000000LL) { rc = -2; break; }
    int a = (int)rnd(rs, n);
    int b = adj[(int64_t)a * degree + rnd(rs, degree)];
    /* c uniformly in the L1 ball of radius L around a, inside the grid */
    int dx = (int)rnd(rs, 2 * L + 1) - L, rem = L - (dx < 0 ? -dx : dx);
    int dy = (int)rnd(rs, 2 * rem + 1) - rem;
    int cx = a / height + dx, cy = a % height + dy;
    if (cx < 0 || cx >= width || cy < 0 || cy >= height) continue;
    int c = cx * height + cy;
    int d = adj[(int64_t)c * degree + rnd(rs, degree)];
    if (grid_dist(b, d, height) > L) continue;
    /* four distinct orbits keep every vertex's degree under the orbit move */
    int oa = orbit_rep(a, width, height, s), ob = orbit_rep(b, width, height, s);
    int oc = orbit_rep(c, width, height, s), od = orbit_rep(d, width, height, s);
    if (oa == ob || oa == oc || oa == od || ob == oc || ob == od || oc == od) continue;
    /* parallel-edge accounting: removing a-b / c-d, adding a-c / b-d */
    int mab = eget(&h, a, b), mcd = eget(&h, c, d);
    int mac = eget(&h, a, c), mbd = eget(&h, b, d);
    int delta = -(mab > 1) - (mcd > 1) + (mac > 0) + (mbd > 0);
    if (delta > 0) continue;
    for (int k = 0; k < orb; k++) {
      int ra = k ? grid_rot(a, width, height, s, k) : a, rb = k ? grid_rot(b, width, height, s, k) : b;
      int rc2 = k ? grid_rot(c, width, height, s, k) : c, rd = k ? grid_rot(d, width, height, s, k) : d;
      if (ODPX_Swap_edges(n, degree, adj, ra, rb, rc2, rd)) { rc = -3; goto out; }
      eadd(&h, ra, rb, -1); eadd(&h, rc2, rd, -1); eadd(&h, ra, rc2, 1); eadd(&h, rb, rd, 1);
    }
    defects += (int64_t)delta * orb;
    moves++;
  }
out:
  free(label); free(queue); eset_free(&h);
  return rc;
}

/* Base table of a symmetric grid in the reference's layout for
 * ODP_Init_aspl_grid_s (utils.c:1116-1151): s = 2 -> rows 0..n/2-1 are the
 * vertices with x < width/2; s = 4 -> row x*(height/2)+y holds vertex
 * x*height+y for x < width/2, y < height/2.  s = 1 copies the table. */
int ODPX_Grid_base_table(int width, int height, int degree, int symmetries, const int *full, int *base)
{
  int n = width * height;
  if (symmetries == 1 || symmetries == 2) {
    memcpy(base, full, sizeof(int) * (size_t)(n / symmetries) * degree);
    return 0;
  }
  if (symmetries != 4) return -1;
  int bh = height / 2;
  for (int x = 0; x < width / 2; x++)
    for (int y = 0; y < bh; y++)
      memcpy(base + (int64_t)(x * bh + y) * degree, full + (int64_t)(x * height + y) * degree,
             sizeof(int) * degree);
  return 0;
}

/* ---- circulant-symmetric general graph ------------------------------------ */

/*
 * Random simple `degree`-regular graph on `nodes` vertices invariant under
 * v -> v + nodes/s (mod nodes), returned as the base table int32[nodes/s][degree]
 * that ODP_Init_aspl_general_s takes (README.md:437-456): row p*(n/s)+m of the
 * full graph is row m shifted by p*(n/s).  Built by pairing base stubs, each
 * pair with a random orbit shift, then repairing loops and repeats by
 * re-pairing with random partners.
 */
int ODPX_Generate_general_s(int nodes, int degree, int symmetries, uint64_t seed, int *base)
{
  int s = symmetries;
  if (s < 1 || nodes % s || degree < 1) return -1;
  int b = nodes / s;
  int64_t nst = (int64_t)b * degree;
  if (nst % 2) return -1;
  int64_t np = nst / 2;
  uint64_t rs[4];
  ODPX_Rng_seed(seed, rs);
  int *st = malloc(sizeof(int) * nst), *sh = malloc(sizeof(int) * np);
  for (int64_t i = 0; i < nst; i++) st[i] = (int)(i / degree);
  for (int64_t i = nst - 1; i > 0; i--) {
    int64_t j = (int64_t)((ODPX_Rng_next(rs) >> 11) % (uint64_t)(i + 1));
    int t = st[i]; st[i] = st[j]; st[j] = t;
  }
  for (int64_t i = 0; i < np; i++) sh[i] = (int)rnd(rs, s);
  int *cnt = malloc(sizeof(int) * b);
  int rc = -2;
  for (int round = 0; round < 200000; round++) {
    /* build rows; detect the first invalid pair */
    for (int i = 0; i < b; i++) cnt[i] = 0;
    int64_t badp = -1;
    for (int64_t p = 0; p < np && badp < 0; p++) {
      int m1 = st[2 * p], m2 = st[2 * p + 1], q = sh[p];
      int t1 = (int)(((int64_t)q * b + m2) % nodes);             /* in row m1 */
      int t2 = (int)(((int64_t)((s - q) % s) * b + m1) % nodes); /* in row m2 */
      if (t1 == m1 || (m1 == m2 && t1 == t2)) { badp = p; break; }
      int *r1 = base + (int64_t)m1 * degree, *r2 = base + (int64_t)m2 * degree;
      for (int j = 0; j < cnt[m1]; j++) if (r1[j] == t1) badp = p;
      for (int j = 0; j < cnt[m2]; j++) if (r2[j] == t2) badp = p;
      if (badp >= 0) break;
      r1[cnt[m1]++] = t1;
      r2[cnt[m2]++] = t2;
    }
    if (badp < 0) { rc = 0; break; }
    /* re-pair the bad pair with a random other pair, fresh shifts */
    int64_t j = (int64_t)((ODPX_Rng_next(rs) >> 11) % (uint64_t)np);
    if (j == badp) continue;
    int t = st[2 * badp + 1]; st[2 * badp + 1] = st[2 * j]; st[2 * j] = t;
    sh[badp] = (int)rnd(rs, s);
    sh[j] = (int)rnd(rs, s);
  }
  free(st); free(sh); free(cnt);
  return rc;
}
