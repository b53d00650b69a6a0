/*
 * odp_mutate.c -- the ODP C API's random graph generation and 2-opt moves
 * (ODP_Srand, ODP_Generate_random_*, ODP_Mutate_adjacency_*,
 * ODP_Restore_adjacency_*): the step before the evaluator in an annealing loop.
 *
 * A caller written against the reference seeds with ODP_Srand, generates a
 * start graph, then mutates / evaluates / restores.  These functions take
 * every random decision the way the reference does -- the same libc rand()
 * draws, mapped to [0, max) as (int)(rand() * max / (RAND_MAX + 1.0)), in the
 * same order, with the same rejections -- so a given seed produces the same
 * graphs and the same move sequence as the reference library, and an
 * annealing run gives the same trajectory.  The move logic is written once for
 * both graph families: a "space" supplies neighbour lookup, storage of an
 * entry, the self-symmetric ("diameter") edge test, the opposite vertex, the
 * edge-length limit and the rotated-edge test of its family.
 *
 * References: utils.c:325-1111 (general), utils.c:380-460, 800-990,
 * 1565-1750 (grid), utils.c:1036-1039 (ODP_Srand).
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdbool.h>
#include "odp.h"
#include "odpx.h"

#define MU_FAIL(...) do { fprintf(stderr, __VA_ARGS__); exit(1); } while (0)
#define GEN_ROUNDS 10      /* mutations per edge when generating (parameter.h:17) */

/* uniform integer in [0, max): from libc rand() (utils.c:356-359), or, inside
 * ODPX_Mutate, from the caller's seeded generator (multiply-shift of its high
 * 32 bits) */
static uint64_t *g_rng = NULL;
static int draw(int max)
{
  if (g_rng) return (int)(((ODPX_Rng_next(g_rng) >> 32) * (uint64_t)max) >> 32);
  return (int)(rand() * ((double)max) / (1.0 + RAND_MAX));
}

static int wrapn(int x, int n)
{
  while (x < 0 || x >= n) x = x < 0 ? x + n : x - n;
  return x;
}

/* grid vertex x*height + y; rotations about the centre (utils.c:67-88) */
static int turn(int v, int width, int height, int s, int angle)
{
  const int x = v / height, y = v % height;
  if (s == 2) {
    if (angle != 180) MU_FAIL("Invalid degree\n");
    return (width - x - 1) * height + (height - y - 1);
  }
  if (s != 4) MU_FAIL("Invalid symmetries(%d)\n", s);
  switch (angle) {
    case 90:  return y * height + (height - x - 1);
    case 180: return (height - x - 1) * height + (height - y - 1);
    case 270: return (height - y - 1) * height + x;
  }
  MU_FAIL("Invalid degree\n");
  return -1;
}

/* row of a first-quadrant (s=4) / left-half (s=2) vertex in the base table (utils.c:213-234) */
static int quadrant_row(int v, int width, int height, int s)
{
  if (s == 1) return v;
  const int x = v / height, y = v % height;
  if (s == 2) {
    if (x >= width / 2) MU_FAIL("Something Wrong ! [id=6] %d %d %d\n", v, height, width / 2);
    return v;
  }
  if (x >= width / 2 || y >= height / 2) MU_FAIL("Something Wrong ! [id=7]\n");
  return x * (height / 2) + y;
}

/* ---- the two families as one "space" -------------------------------------- */

typedef struct {
  bool grid;
  int nodes, degree, s, based;
  int width, height, length;      /* grid only */
  const int *numdeg;              /* per base row, or null */
  int *adj;                       /* [based][degree] */
} Space;

/* the rotation (0, 90, 180, 270) that brings grid vertex v into the base region */
static int grid_home_turn(const Space *S, int v)
{
  if (S->s == 1) return 0;
  const int x = v / S->height, y = v % S->height, bw = S->width / 2, bh = S->height / 2;
  if (S->s == 2) return x < bw ? 0 : 180;
  if (x < bw) return y < bh ? 0 : 270;
  return y >= bh ? 180 : 90;
}

/* base-table row of vertex v */
static int home_row(const Space *S, int v)
{
  if (!S->grid) return v % S->based;
  const int a = grid_home_turn(S, v);
  return quadrant_row(a ? turn(v, S->width, S->height, S->s, a) : v, S->width, S->height, S->s);
}

static int degree_of(const Space *S, int v) { return S->numdeg ? S->numdeg[home_row(S, v)] : S->degree; }

/* j-th neighbour of v in the expanded graph (utils.c:204-209, 237-274) */
static int nbr(const Space *S, int v, int j)
{
  if (!S->grid) return wrapn(S->adj[(v % S->based) * S->degree + j] + (v / S->based) * S->based, S->nodes);
  const int a = grid_home_turn(S, v);
  if (!a) return S->adj[quadrant_row(v, S->width, S->height, S->s) * S->degree + j];
  const int u = S->adj[quadrant_row(turn(v, S->width, S->height, S->s, a), S->width, S->height, S->s) * S->degree + j];
  return turn(u, S->width, S->height, S->s, 360 - a);
}

/* make u the j-th neighbour of v (utils.c:276-281 general, 425-458 grid) */
static void store_grid(int *adj, int width, int height, int degree, int s, int v, int j, int u)
{
  int a;
  if (s == 1) {
    a = 0;
  } else if (s == 2 || s == 4) {
    const int x = v / height, y = v % height, bw = width / 2, bh = s == 2 ? height : height / 2;
    a = s == 2 ? (x < bw ? 0 : 180) : (x < bw ? (y < bh ? 0 : 270) : (y >= bh ? 180 : 90));
  } else {
    return;                              /* no table described: nothing to store */
  }
  if (a) { v = turn(v, width, height, s, a); u = turn(u, width, height, s, a); }
  adj[quadrant_row(v, width, height, s) * degree + j] = u;
}

static void store(Space *S, int v, int j, int u)
{
  if (S->grid) store_grid(S->adj, S->width, S->height, S->degree, S->s, v, j, u);
  else S->adj[(v % S->based) * S->degree + j] = wrapn(u - (v / S->based) * S->based, S->nodes);
}

/* an edge that maps onto itself under the half turn (utils.c:97-110) */
static bool self_symmetric(const Space *S, int u, int v)
{
  if (!S->grid) return S->s % 2 == 0 && abs(u - v) == S->nodes / 2;
  return S->s != 1 && u == turn(v, S->width, S->height, S->s, 180);
}

/* the vertex the half turn takes v to (the general case leaves the shift unreduced) */
static int opposite(const Space *S, int v)
{
  return S->grid ? turn(v, S->width, S->height, S->s, 180) : v + S->nodes / 2;
}

static bool short_enough(const Space *S, int u, int v)
{
  if (!S->grid) return true;
  const int h = S->height;
  return abs(u / h - v / h) + abs(u % h - v % h) <= S->length;
}

/* are edges u0-v0 and u1-v1 images of each other under the symmetry? (utils.c:589-603, 722-740) */
static bool same_orbit(const Space *S, int u0, int v0, int u1, int v1)
{
  if (!S->grid) {
    const int n = S->nodes, b = S->based;
    const int d0 = u0 > v0 ? v0 - u0 + n : v0 - u0;
    const int d1 = u1 > v1 ? v1 - u1 + n : v1 - u1;
    const int d2 = u1 < v1 ? u1 - v1 + n : u1 - v1;
    return (d0 == d1 && u0 % b == u1 % b) || (d0 == d2 && u0 % b == v1 % b);
  }
  if (S->s != 2 && S->s != 4) return false;
  for (int a = 360 / S->s; a < 360; a += 360 / S->s) {
    const int tu = turn(u1, S->width, S->height, S->s, a), tv = turn(v1, S->width, S->height, S->s, a);
    if (u0 == tu && v0 == tv) return true;
  }
  for (int a = 360 / S->s; a < 360; a += 360 / S->s) {
    const int tu = turn(u1, S->width, S->height, S->s, a), tv = turn(v1, S->width, S->height, S->s, a);
    if (u0 == tv && v0 == tu) return true;
  }
  return false;
}

/* position of u in v's row, for the edge u-v found at position u_d of u's row (utils.c:510-555) */
static int position_in(const Space *S, int u, int v, int u_d)
{
  const int d = degree_of(S, v);
  if (u == v) {
    for (int i = 0; i < d; i++)
      if (nbr(S, v, i) == u && i != u_d) return i;
  } else if (self_symmetric(S, u, v)) {
    return u_d;
  } else {
    for (int i = 0; i < d; i++)
      if (nbr(S, v, i) == u) return i;
  }
  if (S->grid) MU_FAIL("Something Wrong ! [id=5]\n");
  MU_FAIL("Something Wrong ! [id=4]\n");
  return -1;
}

/* ---- one move ---------------------------------------------------------------- */

/* rotate edge u-v about u into another image (general: utils.c:571-587) */
static bool turn_edge_general(Space *S, int u, int u_d)
{
  const int n = S->nodes, s = S->s, b = S->based;
  const int v = nbr(S, u, u_d);
  if (self_symmetric(S, u, v)) return false;
  const int v_d = position_in(S, u, v, u_d);
  int nv;
  if ((u - v) % b == 0) {
    if (s <= 4) return false;
    for (;;) {
      nv = v + b * draw(s);
      if (u == nv || v == nv) continue;
      if (wrapn(u - v, n) == wrapn(nv - u, n)) continue;
      if (n % 2 == 1 || (u - nv) % (n / 2) != 0) break;
    }
  } else {
    const int k = s % 2 == 1 ? draw(s - 1) : draw(s);
    nv = k != s - 1 ? v + b * (k + 1) : u + b * (s / 2);
  }
  const int nu = v - (nv - u);
  store(S, u, u_d, nv);
  store(S, v, v_d, nu);
  return true;
}

/* the same for grids (utils.c:790-820) */
static bool turn_edge_grid(Space *S, int u, int u_d)
{
  const int s = S->s, w = S->width, h = S->height;
  const int v = nbr(S, u, u_d);
  if (self_symmetric(S, u, v)) return false;
  if (s == 2 && u == turn(v, w, h, s, 180)) return false;
  if (s == 4 && (u == turn(v, w, h, s, 90) || u == turn(v, w, h, s, 270))) return false;
  const int v_d = position_in(S, u, v, u_d);
  const int k = draw(s);
  if (k != s - 1) {
    const int nv = turn(v, w, h, s, (k + 1) * (360 / s));
    const int nu = turn(u, w, h, s, (s - k - 1) * (360 / s));
    if (!short_enough(S, u, nv)) return false;
    store(S, nv, v_d, u);
    store(S, nu, u_d, v);
  } else {
    const int nv = turn(u, w, h, s, 180), nu = turn(v, w, h, s, 180);
    if (!short_enough(S, u, nv) || !short_enough(S, v, nu)) return false;
    store(S, nv, u_d, u);
    store(S, nu, v_d, v);
  }
  return true;
}

/* one attempt at a 2-opt move; false: rejected, draw again (utils.c:605-730, 822-970) */
static bool try_two_opt(Space *S, ODP_Restore *r)
{
  int u[2], v[2], ud[2], vd[2];
  for (;;) {
    u[0] = draw(S->nodes);
    u[1] = draw(S->nodes);
    if (u[0] == u[1]) continue;
    ud[0] = draw(degree_of(S, u[0]));
    v[0] = nbr(S, u[0], ud[0]);
    if (v[0] == u[1]) continue;
    ud[1] = draw(degree_of(S, u[1]));
    v[1] = nbr(S, u[1], ud[1]);
    if (v[1] == u[0] || v[0] == v[1]) continue;
    break;
  }
  vd[0] = position_in(S, u[0], v[0], ud[0]);
  vd[1] = position_in(S, u[1], v[1], ud[1]);
  if (r)
    for (int i = 0; i < 2; i++) { r->u[i] = u[i]; r->v[i] = v[i]; r->u_d[i] = ud[i]; r->v_d[i] = vd[i]; }

  const bool sym0 = self_symmetric(S, u[0], v[0]), sym1 = self_symmetric(S, u[1], v[1]);
  if (sym0 && sym1) {
    /* both edges self-symmetric: reconnect their endpoints pairwise */
    if (!S->grid && (u[0] - u[1]) % S->based == 0) return false;
    if (S->grid && S->s == 4 &&
        (u[0] == turn(u[1], S->width, S->height, 4, 90) || u[0] == turn(u[1], S->width, S->height, 4, 270)))
      return false;
    int t0, t1;
    if (draw(2)) {
      if (!short_enough(S, u[0], u[1])) return false;
      t0 = u[1]; t1 = u[0];
    } else {
      if (!short_enough(S, u[0], v[1])) return false;
      t0 = v[1]; t1 = v[0];
    }
    store(S, u[0], ud[0], t0);
    store(S, u[1], ud[1], t1);
    return true;
  }
  if (sym0 || sym1) {
    /* one self-symmetric edge (made edge 0) and one ordinary edge */
    if (sym1) {
      int t;
      t = u[0]; u[0] = u[1]; u[1] = t;
      t = ud[0]; ud[0] = ud[1]; ud[1] = t;
      t = v[0]; v[0] = v[1]; v[1] = t;
      t = vd[0]; vd[0] = vd[1]; vd[1] = t;
    }
    const int k = draw(4);
    const int u1o = opposite(S, u[1]), v1o = opposite(S, v[1]);
    int t[4];
    bool ok;
    switch (k) {
      case 0:  ok = short_enough(S, u[0], v[1]) && short_enough(S, u[1], u1o);
               t[0] = v[1]; t[1] = u1o; t[2] = v1o; t[3] = u[0]; break;
      case 1:  ok = short_enough(S, u[0], v1o) && short_enough(S, u[1], u1o);
               t[0] = v1o; t[1] = u1o; t[2] = v[1]; t[3] = v[0]; break;
      case 2:  ok = short_enough(S, u[0], u[1]) && short_enough(S, v[1], v1o);
               t[0] = u[1]; t[1] = u[0]; t[2] = u1o; t[3] = v1o; break;
      default: ok = short_enough(S, u[0], u1o) && short_enough(S, v[1], v1o);
               t[0] = u1o; t[1] = v[0]; t[2] = u[1]; t[3] = v1o; break;
    }
    if (!ok) return false;
    store(S, u[0], ud[0], t[0]);
    store(S, u[1], ud[1], t[1]);
    store(S, v[0], vd[0], t[2]);
    store(S, v[1], vd[1], t[3]);
    return true;
  }
  /* two images of the same edge: turn one of them instead */
  if (same_orbit(S, u[0], v[0], u[1], v[1]))
    return S->grid ? turn_edge_grid(S, u[0], ud[0]) : turn_edge_general(S, u[0], ud[0]);

  int t[4];
  if (draw(2)) {          /* u0-v1 and v0-u1 */
    if (!short_enough(S, u[0], v[1]) || !short_enough(S, v[0], u[1])) return false;
    if (self_symmetric(S, u[0], v[1]) || self_symmetric(S, v[0], u[1])) return false;
    if (same_orbit(S, u[0], v[1], v[0], u[1])) return false;
    t[0] = v[1]; t[1] = v[0]; t[2] = u[1]; t[3] = u[0];
  } else {                /* u0-u1 and v0-v1 */
    if (!short_enough(S, u[0], u[1]) || !short_enough(S, v[0], v[1])) return false;
    if (self_symmetric(S, u[0], u[1]) || self_symmetric(S, v[0], v[1])) return false;
    if (same_orbit(S, u[0], u[1], v[0], v[1])) return false;
    t[0] = u[1]; t[1] = u[0]; t[2] = v[1]; t[3] = v[0];
  }
  store(S, u[0], ud[0], t[0]);
  store(S, u[1], ud[1], t[1]);
  store(S, v[0], vd[0], t[2]);
  store(S, v[1], vd[1], t[3]);
  return true;
}

static void two_opt(Space *S, ODP_Restore *r)
{
  while (!try_two_opt(S, r)) {}
}

/* ---- the table the last symmetric mutation described (read by the restores) --- */

static struct { int nodes, degree, width, height, s; } last;

static void check_parity(int nodes, int degree)
{
  if (nodes % 2 == 1 && degree % 2 == 1)
    MU_FAIL("Nodes(%d) or Degree(%d) must be a multiple of 2.\n", nodes, degree);
}
static void check_div(int nodes, int s)
{
  if (s < 1 || nodes % s != 0) MU_FAIL("nodes(%d) must be divisible by symmetries(%d)\n", nodes, s);
}
static void check_grid_s(int s, int width, int height)
{
  if (s != 1 && s != 2 && s != 4) MU_FAIL("symmetries(%d) must be 1 or 2 or 4\n", s);
  if (s == 2 && width % 2 != 0) MU_FAIL("width(%d) must be divisible by 2\n", width);
  else if (s == 4 && (width % 2 != 0 || height % 2 != 0))
    MU_FAIL("width(%d) and height(%d) must be divisible by 2\n", width, height);
  else if (s == 4 && width != height)
    MU_FAIL("Must be the same as width(%d) and height(%d)\n", width, height);
}

static Space general_space(int nodes, int degree, const int *numdeg, int s, void *adjacency)
{
  Space S = {0};
  S.grid = false; S.nodes = nodes; S.degree = degree; S.s = s; S.based = nodes / s;
  S.numdeg = numdeg; S.adj = (int *)adjacency;
  return S;
}

static Space grid_space(int width, int height, int degree, const int *numdeg, int length, int s, void *adjacency)
{
  Space S = {0};
  S.grid = true; S.nodes = width * height; S.degree = degree; S.s = s; S.based = S.nodes / s;
  S.width = width; S.height = height; S.length = length;
  S.numdeg = numdeg; S.adj = (int *)adjacency;
  return S;
}

/* ---- the API --------------------------------------------------------------- */

/* utils.c:1036-1039 */
void ODP_Srand(unsigned int seed) { srand(seed); }

/* utils.c:752-764 */
void ODP_Mutate_adjacency_general_s(int nodes, int degree, int *num_degrees, int symmetries, ODP_Restore *r,
                                    void *adjacency)
{
  check_div(nodes, symmetries);
  last.nodes = nodes; last.degree = degree; last.s = symmetries;
  Space S = general_space(nodes, degree, num_degrees, symmetries, adjacency);
  two_opt(&S, r);
}

/* utils.c:784-788 */
void ODP_Mutate_adjacency_general(int nodes, int degree, int *num_degrees, ODP_Restore *r, void *adjacency)
{
  ODP_Mutate_adjacency_general_s(nodes, degree, num_degrees, 1, r, adjacency);
}

/* utils.c:973-989 */
void ODP_Mutate_adjacency_grid_s(int width, int height, int degree, int *num_degrees, int length, int symmetries,
                                 ODP_Restore *r, void *adjacency)
{
  check_parity(width * height, degree);
  check_grid_s(symmetries, width, height);
  last.width = width; last.height = height; last.degree = degree; last.s = symmetries;
  Space S = grid_space(width, height, degree, num_degrees, length, symmetries, adjacency);
  two_opt(&S, r);
}

/* utils.c:1565-1574.  Like the reference, this one leaves the table description
 * that ODP_Restore_adjacency_grid reads untouched. */
void ODP_Mutate_adjacency_grid(int width, int height, int degree, int *num_degrees, int length, ODP_Restore *r,
                               void *adjacency)
{
  check_parity(width * height, degree);
  Space S = grid_space(width, height, degree, num_degrees, length, 1, adjacency);
  two_opt(&S, r);
}

/* One 2-opt move of the reference's kind (the move of ODP_Mutate_adjacency_*,
 * utils.c:571-990) drawn from a seeded generator instead of libc rand();
 * family 0 general (nodes_or_width = nodes), 1 grid (width, height, length).
 * It records the table for ODP_Restore_adjacency_general / _grid in every
 * case.  Returns 0, or -1 for invalid parameters. */
int ODPX_Mutate(int family, int nodes_or_width, int height, int degree, int length, int symmetries,
                void *adjacency, uint64_t rng[4], ODP_Restore *r)
{
  if (!adjacency || !rng || !r || degree < 1 || symmetries < 1) return -1;
  Space S;
  if (family == 0) {
    if (nodes_or_width % symmetries != 0) return -1;
    last.nodes = nodes_or_width; last.degree = degree; last.s = symmetries;
    S = general_space(nodes_or_width, degree, NULL, symmetries, adjacency);
  } else {
    const int w = nodes_or_width;
    if (symmetries != 1 && symmetries != 2 && symmetries != 4) return -1;
    if ((symmetries == 2 && w % 2) || (symmetries == 4 && (w % 2 || height % 2 || w != height))) return -1;
    last.width = w; last.height = height; last.degree = degree; last.s = symmetries;
    S = grid_space(w, height, degree, NULL, length, symmetries, adjacency);
  }
  g_rng = rng;
  two_opt(&S, r);
  g_rng = NULL;
  return 0;
}

/* utils.c:500-507: puts back the two edges a general mutation replaced */
void ODP_Restore_adjacency_general(ODP_Restore r, void *adjacency)
{
  if (last.s < 1) MU_FAIL("ODP_Restore_adjacency_general: no earlier ODP_Mutate_adjacency_general[_s]\n");
  int *adj = (int *)adjacency;
  const int b = last.nodes / last.s;
  for (int i = 0; i < 2; i++) {
    adj[(r.u[i] % b) * last.degree + r.u_d[i]] = wrapn(r.v[i] - (r.u[i] / b) * b, last.nodes);
    adj[(r.v[i] % b) * last.degree + r.v_d[i]] = wrapn(r.u[i] - (r.v[i] / b) * b, last.nodes);
  }
}

/* utils.c:461-467: the same for grids, with the table that the last
 * ODP_Mutate_adjacency_grid_s described (none: nothing is stored) */
void ODP_Restore_adjacency_grid(ODP_Restore r, void *adjacency)
{
  for (int i = 0; i < 2; i++) {
    store_grid((int *)adjacency, last.width, last.height, last.degree, last.s, r.u[i], r.u_d[i], r.v[i]);
    store_grid((int *)adjacency, last.width, last.height, last.degree, last.s, r.v[i], r.v_d[i], r.u[i]);
  }
}

/* ---- generation ------------------------------------------------------------- */

/* true when some vertex is unreachable from vertex 0 (utils.c:283-347) */
static bool disconnected(const Space *S)
{
  const int n = S->nodes;
  char *seen = calloc((size_t)n, 1);
  int *queue = malloc(sizeof(int) * (size_t)n);
  int head = 0, tail = 0, reached = 1;
  queue[tail++] = 0;
  seen[0] = 1;
  while (head < tail) {
    const int v = queue[head++];
    for (int j = 0; j < S->degree; j++) {
      int u;
      if (S->grid && S->s != 1) u = nbr(S, v, j);
      else {
        u = S->adj[(v % S->based) * S->degree + j] + (v / S->based) * S->based;
        if (u >= n) u -= n;
      }
      if (!seen[u]) { seen[u] = 1; reached++; queue[tail++] = u; }
    }
  }
  free(seen);
  free(queue);
  return reached != n;
}

/* a ring with degree/2 parallel edges between neighbours, plus a perfect
 * matching across the ring for odd degrees (utils.c:1041-1063) */
static void ring_graph(int nodes, int degree, int (*edge)[2])
{
  const int half = degree / 2;
  for (int i = 0; i < nodes; i++)
    for (int j = 0; j < half; j++) {
      edge[i * half + j][0] = i;
      edge[i * half + j][1] = i + 1 < nodes ? i + 1 : 0;
    }
  if (degree % 2 == 1)
    for (int i = 0; i < nodes / 2; i++) {
      edge[half * nodes + i][0] = i;
      edge[half * nodes + i][1] = i + nodes / 2;
    }
}

/* utils.c:1065-1084 */
void ODP_Generate_random_general(int nodes, int degree, int (*edge)[2])
{
  check_parity(nodes, degree);
  ring_graph(nodes, degree, edge);
  const int lines = nodes * degree / 2;
  int *adj = malloc(sizeof(int) * (size_t)nodes * degree);
  ODP_Conv_edge2adjacency_general(nodes, lines, degree, edge, adj);
  for (long i = 0; i < (long)lines * GEN_ROUNDS; i++) ODP_Mutate_adjacency_general(nodes, degree, NULL, NULL, adj);
  Space S = general_space(nodes, degree, NULL, 1, adj);
  while (disconnected(&S)) ODP_Mutate_adjacency_general(nodes, degree, NULL, NULL, adj);
  ODP_Conv_adjacency2edge_general(nodes, degree, NULL, adj, edge);
  free(adj);
}

/* utils.c:1086-1114 */
void ODP_Generate_random_general_s(int nodes, int degree, int symmetries, int (*edge)[2])
{
  check_div(nodes, symmetries);
  check_parity(nodes, degree);
  if (symmetries == 1) { ODP_Generate_random_general(nodes, degree, edge); return; }
  ring_graph(nodes, degree, edge);
  const int based = nodes / symmetries, lines = nodes * degree / 2;
  int *adj = malloc(sizeof(int) * (size_t)based * degree);
  ODP_Conv_edge2adjacency_general_s(nodes, lines, degree, edge, symmetries, adj);
  for (long i = 0; i < (long)lines * GEN_ROUNDS; i++)
    ODP_Mutate_adjacency_general_s(nodes, degree, NULL, symmetries, NULL, adj);
  Space S = general_space(nodes, degree, NULL, symmetries, adj);
  while (disconnected(&S)) ODP_Mutate_adjacency_general_s(nodes, degree, NULL, symmetries, NULL, adj);
  ODP_Conv_adjacency2edge_general_s(nodes, degree, NULL, adj, symmetries, edge);
  free(adj);
}

/* the lattice start graph of create-lattice.c (research.nii.ac.jp/graphgolf):
 * degree parallel edges between columns 2x and 2x+1, an odd last column
 * paired vertically, and self-loops on the last vertex of an odd grid
 * (utils.c:1576-1611) */
static void lattice_graph(int width, int height, int degree, int (*edge)[2])
{
  int i = 0;
  for (int x = 0; x < width / 2; x++)
    for (int y = 0; y < height; y++)
      for (int k = 0; k < degree; k++, i++) {
        edge[i][0] = y + 2 * x * height;
        edge[i][1] = edge[i][0] + height;
      }
  if (width % 2 == 1) {
    for (int y = 0; y < height / 2; y++)
      for (int k = 0; k < degree; k++, i++) {
        edge[i][0] = (width - 1) * height + 2 * y;
        edge[i][1] = edge[i][0] + 1;
      }
    if (height % 2 == 1)
      for (int k = 0; k < degree / 2; k++, i++) edge[i][0] = edge[i][1] = width * height - 1;
  }
}

/* utils.c:1577-1625 */
void ODP_Generate_random_grid(int width, int height, int degree, int length, int (*edge)[2])
{
  const int nodes = width * height;
  check_parity(nodes, degree);
  lattice_graph(width, height, degree, edge);
  const int lines = nodes * degree / 2;
  int *adj = malloc(sizeof(int) * (size_t)nodes * degree);
  ODP_Conv_edge2adjacency_grid(width, height, lines, degree, edge, adj);
  for (long i = 0; i < (long)lines * GEN_ROUNDS; i++)
    ODP_Mutate_adjacency_grid(width, height, degree, NULL, length, NULL, adj);
  Space S = general_space(nodes, degree, NULL, 1, adj);
  while (disconnected(&S)) ODP_Mutate_adjacency_grid(width, height, degree, NULL, length, NULL, adj);
  ODP_Conv_adjacency2edge_grid(width, height, degree, NULL, adj, edge);
  free(adj);
}

/* copies edges [0, cnt) into the other symmetry images: s = 2 the half turn
 * at cnt + i, s = 4 the quarter turns at cnt + i, 2 cnt + i, 3 cnt + i */
static void spread_images(int (*edge)[2], int cnt, int width, int height, int s)
{
  for (int i = 0; i < cnt; i++)
    for (int c = 1; c < s; c++)
      for (int j = 0; j < 2; j++) edge[c * cnt + i][j] = turn(edge[i][j], width, height, s, c * (360 / s));
}

/* edge k and its three quarter-turn images at k+1..k+3 */
static void quad_edge(int (*edge)[2], int k, int a, int b, int width, int height)
{
  edge[k][0] = a;
  edge[k][1] = b;
  for (int c = 1; c < 4; c++) {
    edge[k + c][0] = turn(a, width, height, 4, 90 * c);
    edge[k + c][1] = turn(b, width, height, 4, 90 * c);
  }
}

/* utils.c:1627-1750 */
void ODP_Generate_random_grid_s(int width, int height, int degree, int length, int symmetries, int (*edge)[2])
{
  check_parity(width * height, degree);
  check_grid_s(symmetries, width, height);
  if (symmetries == 1) { ODP_Generate_random_grid(width, height, degree, length, edge); return; }
  const int nodes = width * height, lines = nodes * degree / 2, based = nodes / symmetries;
  const int based_lines = lines / symmetries;
  const int bw = width / 2, bh = symmetries == 2 ? height : height / 2;
  /* a random base region (its own lattice start and mutations), moved into
   * place and copied into the other images */
  const bool whole = based % 2 == 0 || degree % 2 == 0;
  const int bdeg = whole ? degree : degree - 1;
  ODP_Generate_random_grid(bw, bh, bdeg, length, edge);
  const int cnt = whole ? based_lines : based * (degree - 1) / 2;
  for (int i = 0; i < cnt; i++)
    for (int j = 0; j < 2; j++) edge[i][j] = (edge[i][j] / bh) * height + edge[i][j] % bh;
  spread_images(edge, cnt, width, height, symmetries);
  if (!whole) {
    /* odd base regions: one more edge per vertex, pairing across the regions */
    int k = cnt * symmetries;
    if (symmetries == 2) {
      for (int x = 0; x < width; x += 2)
        for (int y = 0; y < height; y++, k++) {
          edge[k][0] = y + x * height;
          edge[k][1] = edge[k][0] + height;
        }
    } else {
      for (int x = 0; x < bw / 2; x++)
        for (int y = 0; y < bh; y++, k += 4) quad_edge(edge, k, 2 * x * height + y, 2 * x * height + y + height, width, height);
      if (bw % 2 == 1) {
        for (int y = 0; y < bh / 2; y++, k += 4)
          quad_edge(edge, k, (bw - 1) * height + 2 * y, (bw - 1) * height + 2 * y + 1, width, height);
        if (bh % 2 == 1) {
          if (length <= 1) MU_FAIL("Graphs with odd width/2 and length/2 and length 1 cannot be created.");
          edge[k][0] = (bw - 1) * height + (bh - 1);
          edge[k][1] = bw * height + bh;
          edge[k + 1][0] = (bw - 1) * height + bh;
          edge[k + 1][1] = bw * height + (bh - 1);
        }
      }
    }
  }
  int *adj = malloc(sizeof(int) * (size_t)based * degree);
  ODP_Conv_edge2adjacency_grid_s(width, height, lines, degree, edge, symmetries, adj);
  for (long i = 0; i < (long)based_lines * GEN_ROUNDS; i++)
    ODP_Mutate_adjacency_grid_s(width, height, degree, NULL, length, symmetries, NULL, adj);
  Space S = grid_space(width, height, degree, NULL, length, symmetries, adj);
  while (disconnected(&S)) ODP_Mutate_adjacency_grid_s(width, height, degree, NULL, length, symmetries, NULL, adj);
  ODP_Conv_adjacency2edge_grid_s(width, height, degree, NULL, adj, symmetries, edge);
  free(adj);
}
