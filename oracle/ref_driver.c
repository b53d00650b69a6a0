/*
 * ref_driver.c -- a small harness that runs the REFERENCE library
 * (aspl.c + utils.c from the reference sources, compiled by oracle/Makefile
 * into oracle/_ref/) on one graph.  TEST INFRASTRUCTURE ONLY: used to pin the
 * oracle restatement and as the CPU baseline; never part of the product.
 *
 * Input: a binary graph file (little-endian int32):
 *   header[8] = { 0x3150444F ("ODP1"), family (0 general, 1 grid),
 *                 nodes_or_width, height, degree, symmetries, has_num_degrees, repeat }
 *   num_degrees[(s == 1) ? nodes : nodes/s]   if has_num_degrees
 *   adjacency[(nodes/s) * degree]
 * Output (stdout): "<diameter> <sum> <ASPL as %a>" then, when repeat > 0,
 *   one line "t <seconds>" per timed ODP_Set_aspl call (after one warm-up).
 * The algorithm kind follows ODP_ASPL in the environment, as in the reference
 * (utils.c:1238-1266).
 */
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>
#include "odp.h"

static double now(void)
{
  struct timeval t;
  gettimeofday(&t, NULL);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

int main(int argc, char **argv)
{
  if (argc != 2) { fprintf(stderr, "usage: %s graph.bin\n", argv[0]); return 2; }
  FILE *fp = fopen(argv[1], "rb");
  if (!fp) { fprintf(stderr, "cannot open %s\n", argv[1]); return 2; }
  int h[8];
  if (fread(h, sizeof(int), 8, fp) != 8 || h[0] != 0x3150444F) { fprintf(stderr, "bad header\n"); return 2; }
  int family = h[1], nw = h[2], height = h[3], degree = h[4], s = h[5], has_nd = h[6], repeat = h[7];
  int nodes = family ? nw * height : nw;
  int based = nodes / s;
  int *nd = NULL;
  if (has_nd) {
    int cnt = (s == 1) ? nodes : based;
    nd = malloc(sizeof(int) * cnt);
    if (fread(nd, sizeof(int), cnt, fp) != (size_t)cnt) { fprintf(stderr, "short num_degrees\n"); return 2; }
  }
  size_t na = (size_t)based * degree;
  int *adj = malloc(sizeof(int) * na);
  if (fread(adj, sizeof(int), na, fp) != na) { fprintf(stderr, "short adjacency\n"); return 2; }
  fclose(fp);

  if (!family) {
    if (s == 1) ODP_Init_aspl_general(nodes, degree, nd);
    else        ODP_Init_aspl_general_s(nodes, degree, nd, s);
  } else {
    if (s == 1) ODP_Init_aspl_grid(nw, height, degree, nd);
    else        ODP_Init_aspl_grid_s(nw, height, degree, nd, s);
  }
  int diameter;
  long sum;
  double aspl;
  ODP_Set_aspl(adj, &diameter, &sum, &aspl);
  printf("%d %ld %a\n", diameter, sum, aspl);
  for (int r = 0; r < repeat; r++) {
    double t0 = now();
    ODP_Set_aspl(adj, &diameter, &sum, &aspl);
    printf("t %.9f\n", now() - t0);
    fflush(stdout);
  }
  ODP_Finalize_aspl();
  free(adj);
  free(nd);
  return 0;
}
