/*
 * odp_utils.c -- host-side graph utilities of the ODP C API (edge I/O,
 * edge <-> adjacency conversion, degree/size queries, lower bounds).
 * They are not on the evaluator's hot path; the library exports them so that
 * programs written against the reference (its sample and annealing programs) link
 * against libodp_hip unchanged.  Each function follows the reference
 * definition cited above it, including its error behaviour (print, exit(1)).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>
#include <math.h>
#include "odp.h"

#define ODP_FAIL(...) do { fprintf(stderr, __VA_ARGS__); exit(1); } while (0)
#define NOT_DEFINED (-1)

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* x = w*height + h  (common.h:34-35) */
static int col_of(int v, int height) { return v / height; }
static int row_of(int v, int height) { return v % height; }

static int manhattan(int u, int v, int height)
{
  return abs(col_of(u, height) - col_of(v, height)) + abs(row_of(u, height) - row_of(v, height));
}

/* utils.c:67-88 */
static int rotate_vertex(int v, int width, int height, int symmetries, int angle)
{
  int w = col_of(v, height), h = row_of(v, height);
  if (symmetries == 2) return (width - w - 1) * height + (height - h - 1);
  if (angle == 90)  return h * height + (height - w - 1);
  if (angle == 180) return (height - w - 1) * height + (height - h - 1);
  return (height - h - 1) * height + w;
}

/* utils.c:113-133 */
static void check_symmetries(int nodes, int symmetries)
{
  if (symmetries < 1 || nodes % symmetries != 0)
    ODP_FAIL("nodes(%d) must be divisible by symmetries(%d)\n", nodes, symmetries);
}
static void check_symmetries_grid(int symmetries)
{
  if (symmetries != 1 && symmetries != 2 && symmetries != 4)
    ODP_FAIL("symmetries(%d) must be 1 or 2 or 4\n", symmetries);
}
static void check_symmetries_wh(int symmetries, int width, int height)
{
  if (symmetries == 2 && width % 2 != 0) ODP_FAIL("width(%d) must be divisible by 2\n", width);
  if (symmetries == 4 && (width % 2 != 0 || height % 2 != 0))
    ODP_FAIL("width(%d) and height(%d) must be divisible by 2\n", width, height);
  if (symmetries == 4 && width != height)
    ODP_FAIL("Must be the same as width(%d) and height(%d)\n", width, height);
}

/* utils.c:195-201 */
static int wrap(int x, int nodes)
{
  while (x < 0 || x >= nodes) x = (x < 0) ? x + nodes : x - nodes;
  return x;
}

/* ---- printing ------------------------------------------------------------ */

void ODP_Print_adjacency(int nodes, int degree, int *num_degrees, void *adjacency)
{
  const int *adj = (const int *)adjacency;
  for (int i = 0; i < nodes; i++) {
    printf("%d ", i);
    int d = num_degrees ? num_degrees[i] : degree;
    for (int j = 0; j < d; j++)
      if (adj[(long)i * degree + j] != NOT_DEFINED) printf("%3d", adj[(long)i * degree + j]);
    printf("\n");
  }
}

void ODP_Print_edge_general(int lines, int (*edge)[2])
{
  for (int i = 0; i < lines; i++) printf("%d %d\n", edge[i][0], edge[i][1]);
}

void ODP_Print_edge_grid(int lines, int height, int (*edge)[2])
{
  for (int i = 0; i < lines; i++)
    printf("%d,%d %d,%d\n", col_of(edge[i][0], height), row_of(edge[i][0], height),
           col_of(edge[i][1], height), row_of(edge[i][1], height));
}

/* ---- checks and queries --------------------------------------------------- */

bool ODP_Check_loop(int lines, int (*edge)[2])
{
  for (int i = 0; i < lines; i++)
    if (edge[i][0] == edge[i][1]) {
      printf("Loop is found in line %d\n", i + 1);
      return true;
    }
  return false;
}

bool ODP_Check_multiple_edges(int lines, int (*edge)[2])
{
  for (int i = 0; i < lines; i++)
    for (int j = i + 1; j < lines; j++) {
      int a = edge[i][0], b = edge[i][1], c = edge[j][0], d = edge[j][1];
      if ((a == c && b == d) || (a == d && b == c)) {
        printf("Multiple edeges are found in lines %d %d\n", i + 1, j + 1);
        return true;
      }
    }
  return false;
}

int ODP_Get_length(int lines, int height, int (*edge)[2])
{
  int length = 0;
  for (int i = 0; i < lines; i++) length = imax(length, manhattan(edge[i][0], edge[i][1], height));
  return length;
}

bool ODP_Check_general(char *fname)
{
  FILE *fp = fopen(fname, "r");
  if (!fp) ODP_FAIL("File not found\n");
  int a = 1, b = -1;
  if (fscanf(fp, "%d %d", &a, &b) < 0) b = -1;
  fclose(fp);
  return b != -1;
}

int ODP_Get_degree(int nodes, int lines, int (*edge)[2])
{
  int *cnt = calloc((size_t)nodes, sizeof(int));
  for (int i = 0; i < lines; i++) { cnt[edge[i][0]]++; cnt[edge[i][1]]++; }
  int degree = cnt[0];
  for (int i = 1; i < nodes; i++) degree = imax(degree, cnt[i]);
  free(cnt);
  return degree;
}

int ODP_Get_nodes(int lines, int (*edge)[2])
{
  int m = 0;
  for (int i = 0; i < lines; i++) m = imax(m, imax(edge[i][0], edge[i][1]));
  return m + 1;
}

int ODP_Get_lines(char *fname)
{
  FILE *fp = fopen(fname, "r");
  if (!fp) ODP_FAIL("File not found\n");
  int lines = 0, c;
  while ((c = fgetc(fp)) != EOF)
    if (c == '\n') lines++;
  fclose(fp);
  return lines;
}

void ODP_Set_degrees(int nodes, int lines, int (*edge)[2], int *num_degrees)
{
  for (int i = 0; i < nodes; i++) num_degrees[i] = 0;
  for (int i = 0; i < lines; i++) { num_degrees[edge[i][0]]++; num_degrees[edge[i][1]]++; }
}

/* ---- edge files ------------------------------------------------------------ */

void ODP_Read_edge_general(char *fname, int (*edge)[2])
{
  FILE *fp = fopen(fname, "r");
  if (!fp) ODP_FAIL("File not found\n");
  int a, b, i = 0;
  while (fscanf(fp, "%d %d", &a, &b) != EOF) { edge[i][0] = a; edge[i][1] = b; i++; }
  fclose(fp);
}

void ODP_Read_edge_grid(char *fname, int *w, int *h, int (*edge)[2])
{
  FILE *fp = fopen(fname, "r");
  if (!fp) ODP_FAIL("File not found\n");
  int x0, y0, x1, y1;
  *w = 0; *h = 0;
  while (fscanf(fp, "%d,%d %d,%d", &x0, &y0, &x1, &y1) != EOF) {
    *w = imax(*w, imax(x0, x1));
    *h = imax(*h, imax(y0, y1));
  }
  *w += 1; *h += 1;
  rewind(fp);
  int i = 0;
  while (fscanf(fp, "%d,%d %d,%d", &x0, &y0, &x1, &y1) != EOF) {
    edge[i][0] = x0 * (*h) + y0;
    edge[i][1] = x1 * (*h) + y1;
    i++;
  }
  fclose(fp);
}

void ODP_Write_edge_general(int lines, int (*edge)[2], char *fname)
{
  FILE *fp = fopen(fname, "w");
  if (!fp) ODP_FAIL("Cannot open %s\n", fname);
  for (int i = 0; i < lines; i++) fprintf(fp, "%d %d\n", edge[i][0], edge[i][1]);
  fclose(fp);
}

void ODP_Write_edge_grid(int lines, int height, int (*edge)[2], char *fname)
{
  FILE *fp = fopen(fname, "w");
  if (!fp) ODP_FAIL("Cannot open %s\n", fname);
  for (int i = 0; i < lines; i++)
    fprintf(fp, "%d,%d %d,%d\n", col_of(edge[i][0], height), row_of(edge[i][0], height),
            col_of(edge[i][1], height), row_of(edge[i][1], height));
  fclose(fp);
}

/* ---- conversions ------------------------------------------------------------ */

/* utils.c:1012-1034: only base vertices (id < nodes/s) get rows; unused slots
 * stay NOT_DEFINED. */
void ODP_Conv_edge2adjacency_general_s(int nodes, int lines, int degree, int (*edge)[2], int symmetries,
                                       void *adjacency)
{
  check_symmetries(nodes, symmetries);
  int *adj = (int *)adjacency;
  int based = nodes / symmetries;
  int *fill = calloc((size_t)based, sizeof(int));
  for (long i = 0; i < (long)based * degree; i++) adj[i] = NOT_DEFINED;
  for (int i = 0; i < lines; i++) {
    int a = edge[i][0], b = edge[i][1];
    if (a < based) adj[(long)a * degree + fill[a]++] = b;
    if (b < based) adj[(long)b * degree + fill[b]++] = a;
  }
  free(fill);
}

void ODP_Conv_edge2adjacency_general(int nodes, int lines, int degree, int (*edge)[2], void *adjacency)
{
  ODP_Conv_edge2adjacency_general_s(nodes, lines, degree, edge, 1, adjacency);
}

void ODP_Conv_edge2adjacency_grid(int width, int height, int lines, int degree, int (*edge)[2], void *adjacency)
{
  ODP_Conv_edge2adjacency_general_s(width * height, lines, degree, edge, 1, adjacency);
}

/* first-quadrant local index (utils.c:213-234) */
static int local_index_grid(int x, int width, int height, int symmetries)
{
  if (symmetries != 4) return x;
  (void)width;
  return col_of(x, height) * (height / 2) + row_of(x, height);
}

/* utils.c:1116-1151 */
void ODP_Conv_edge2adjacency_grid_s(int width, int height, int lines, int degree, int (*edge)[2], int symmetries,
                                    void *adjacency)
{
  check_symmetries_grid(symmetries);
  check_symmetries_wh(symmetries, width, height);
  if (symmetries == 1) { ODP_Conv_edge2adjacency_grid(width, height, lines, degree, edge, adjacency); return; }
  int *adj = (int *)adjacency;
  int nodes = width * height, based = nodes / symmetries;
  int *fill = calloc((size_t)based, sizeof(int));
  for (long i = 0; i < (long)based * degree; i++) adj[i] = NOT_DEFINED;
  int bw = width / 2, bh = (symmetries == 2) ? height : height / 2;
  for (int i = 0; i < lines; i++) {
    int a = edge[i][0], b = edge[i][1];
    if (col_of(a, height) < bw && row_of(a, height) < bh) {
      int la = local_index_grid(a, width, height, symmetries);
      adj[(long)la * degree + fill[la]++] = b;
    }
    if (col_of(b, height) < bw && row_of(b, height) < bh) {
      int lb = local_index_grid(b, width, height, symmetries);
      adj[(long)lb * degree + fill[lb]++] = a;
    }
  }
  free(fill);
}

static int count_lines(int nodes, int degree, const int *num_degrees)
{
  if (!num_degrees) return (nodes * degree) / 2;
  int lines = 0;
  for (int i = 0; i < nodes; i++) lines += num_degrees[i];
  if (lines % 2 == 1) ODP_FAIL("Something Wrong ! [id=2]\n");
  return lines / 2;
}

/* utils.c:1506-1533: each edge once (i < v); a self-loop appears twice in
 * its row and is emitted once per pair. */
void ODP_Conv_adjacency2edge_general(int nodes, int degree, int *num_degrees, void *adjacency, int (*edge)[2])
{
  const int *adj = (const int *)adjacency;
  int lines = 0, loops = 0;
  for (int i = 0; i < nodes; i++) {
    int d = num_degrees ? num_degrees[i] : degree;
    for (int j = 0; j < d; j++) {
      int v = adj[(long)i * degree + j];
      if (i < v) { edge[lines][0] = i; edge[lines][1] = v; lines++; }
      else if (i == v) {
        loops++;
        if (loops % 2 == 0) { edge[lines][0] = i; edge[lines][1] = v; lines++; }
      }
    }
  }
  if (loops % 2 == 1 || lines != count_lines(nodes, degree, num_degrees))
    ODP_FAIL("Something Wrong ! [id=1]\n");
}

void ODP_Conv_adjacency2edge_grid(int width, int height, int degree, int *num_degrees, void *adjacency,
                                  int (*edge)[2])
{
  ODP_Conv_adjacency2edge_general(width * height, degree, num_degrees, adjacency, edge);
}

/* utils.c:991-1010: the base rows shifted into every orbit, then as above */
void ODP_Conv_adjacency2edge_general_s(int nodes, int degree, int *num_degrees, void *adjacency, int symmetries,
                                       int (*edge)[2])
{
  check_symmetries(nodes, symmetries);
  const int *adj = (const int *)adjacency;
  int based = nodes / symmetries;
  int *full = malloc(sizeof(int) * (size_t)nodes * degree);
  for (int p = 0; p < symmetries; p++)
    for (int j = 0; j < based; j++) {
      int d = num_degrees ? num_degrees[j] : degree;
      for (int k = 0; k < d; k++)
        full[(long)(p * based + j) * degree + k] = wrap(adj[(long)j * degree + k] + p * based, nodes);
    }
  ODP_Conv_adjacency2edge_general(nodes, degree, num_degrees, full, edge);
  free(full);
}

/* utils.c:152-193: rotate the first-quadrant (s=4) or left-half (s=2) rows */
void ODP_Conv_adjacency2edge_grid_s(int width, int height, int degree, int *num_degrees, void *adjacency,
                                    int symmetries, int (*edge)[2])
{
  check_symmetries_grid(symmetries);
  check_symmetries_wh(symmetries, width, height);
  if (symmetries == 1) { ODP_Conv_adjacency2edge_grid(width, height, degree, num_degrees, adjacency, edge); return; }
  const int *adj = (const int *)adjacency;
  int nodes = width * height, based = nodes / symmetries;
  int *full = malloc(sizeof(int) * (size_t)nodes * degree);
  if (symmetries == 2) {
    for (int i = 0; i < based; i++) {
      int d = num_degrees ? num_degrees[i] : degree;
      int ri = rotate_vertex(i, width, height, 2, 180);
      for (int k = 0; k < d; k++) {
        int v = adj[(long)i * degree + k];
        full[(long)i * degree + k] = v;
        full[(long)ri * degree + k] = rotate_vertex(v, width, height, 2, 180);
      }
    }
  } else {
    int bh = height / 2;
    for (int i = 0; i < based; i++) {
      int d = num_degrees ? num_degrees[i] : degree;
      int j = (i / bh) * height + (i % bh);
      for (int k = 0; k < d; k++) {
        int v = adj[(long)i * degree + k];
        full[(long)j * degree + k] = v;
        for (int a = 90; a <= 270; a += 90)
          full[(long)rotate_vertex(j, width, height, 4, a) * degree + k] = rotate_vertex(v, width, height, 4, a);
      }
    }
  }
  ODP_Conv_adjacency2edge_grid(width, height, degree, num_degrees, full, edge);
  free(full);
}

/* ---- lower bounds ------------------------------------------------------------ */

/* utils.c:1432-1453: Moore bound of a degree-regular graph */
void ODP_Set_lbounds_general(int nodes, int degree, int *low_diameter, double *low_ASPL)
{
  int diam = -1, reached = 1, r = 1;
  double aspl = 0.0;
  for (;;) {
    int next = reached + degree * pow(degree - 1, r - 1);
    if (next >= nodes) break;
    reached = next;
    aspl += r * degree * pow(degree - 1, r - 1);
    diam = r++;
  }
  diam++;
  aspl += diam * (nodes - reached);
  aspl /= (nodes - 1);
  *low_diameter = diam;
  *low_ASPL = aspl;
}

/* utils.c:1456-1504: per-source lattice-distance histogram capped by the
 * Moore bound, for edges of Manhattan length <= length */
void ODP_Set_lbounds_grid(int m, int n, int degree, int length, int *low_diameter, double *low_ASPL)
{
  int mn = m * n, current = degree, ii;
  size_t cap = (size_t)imax(mn, m + n) + 8;
  int *moore = malloc(sizeof(int) * cap), *hist = malloc(sizeof(int) * cap), *mh = malloc(sizeof(int) * cap);
  double sum = 0;
  moore[0] = 1;
  moore[1] = degree + 1;
  for (ii = 2;; ii++) {
    current = current * (degree - 1);
    moore[ii] = moore[ii - 1] + current;
    if (moore[ii] >= mn || current <= 0) { moore[ii] = mn; break; }   /* degree <= 1 never grows */
  }
  int maxhop = imax((m + n - 2 + (length - 1)) / length, ii);
  for (int i = ii + 1; i <= maxhop; i++) moore[i] = mn;
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) {
      for (int k = 0; k <= maxhop; k++) hist[k] = 0;
      for (int i2 = 0; i2 < m; i2++)
        for (int j2 = 0; j2 < n; j2++) hist[(abs(i - i2) + abs(j - j2) + length - 1) / length]++;
      for (int k = 1; k <= maxhop; k++) hist[k] += hist[k - 1];
      for (int k = 0; k <= maxhop; k++) mh[k] = imin(hist[k], moore[k]);
      for (int k = 1; k <= maxhop; k++) sum += (double)(mh[k] - mh[k - 1]) * k;
    }
  int dboth = 0;
  while (mh[dboth] != mn) dboth++;
  *low_diameter = dboth;
  *low_ASPL = sum / ((double)mn * (mn - 1));
  free(moore); free(hist); free(mh);
}
