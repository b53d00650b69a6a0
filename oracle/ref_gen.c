/*
 * ref_gen.c -- runs the REFERENCE library's random generation and 2-opt moves
 * (utils.c from the reference sources, compiled by oracle/Makefile into
 * oracle/_ref/).  TEST INFRASTRUCTURE ONLY: pins libodp_hip's ODP_Srand /
 * ODP_Generate_random_* / ODP_Mutate_adjacency_* / ODP_Restore_adjacency_*.
 *
 *   odp_ref_gen general N D S SEED MOVES
 *   odp_ref_gen grid    W H D LENGTH S SEED MOVES [grid_s_for_s1]
 *
 * Seeds with ODP_Srand, generates a graph, prints its edges, converts it to the
 * (base) adjacency table, then makes MOVES moves, restoring every third one,
 * and prints each move's undo record and the table's checksum, and the final
 * table.  For a grid with S = 1 the moves use ODP_Mutate_adjacency_grid unless
 * the last argument is 1 (then ODP_Mutate_adjacency_grid_s with S = 1).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "odp.h"

static unsigned long long fnv(const int *a, long n)
{
  unsigned long long h = 1469598103934665603ULL;
  for (long i = 0; i < n; i++) {
    unsigned int x = (unsigned int)a[i];
    for (int b = 0; b < 4; b++) { h ^= (x >> (8 * b)) & 0xff; h *= 1099511628211ULL; }
  }
  return h;
}

int main(int argc, char **argv)
{
  if (argc < 2) { fprintf(stderr, "usage: see source\n"); return 2; }
  const int grid = strcmp(argv[1], "grid") == 0;
  if ((!grid && argc != 7) || (grid && argc != 9 && argc != 10)) { fprintf(stderr, "bad arguments\n"); return 2; }
  int w = 0, h = 0, nodes, d, len = 0, s, moves, use_s = 0;
  unsigned seed;
  if (!grid) {
    nodes = atoi(argv[2]); d = atoi(argv[3]); s = atoi(argv[4]); seed = (unsigned)atoi(argv[5]); moves = atoi(argv[6]);
  } else {
    w = atoi(argv[2]); h = atoi(argv[3]); d = atoi(argv[4]); len = atoi(argv[5]); s = atoi(argv[6]);
    seed = (unsigned)atoi(argv[7]); moves = atoi(argv[8]); use_s = argc == 10 ? atoi(argv[9]) : 0;
    nodes = w * h;
  }
  const int lines = nodes * d / 2, rows = nodes / s;
  int (*edge)[2] = malloc(sizeof(int) * 2 * lines);
  int *adj = malloc(sizeof(int) * rows * d);
  ODP_Srand(seed);
  if (!grid) ODP_Generate_random_general_s(nodes, d, s, edge);
  else ODP_Generate_random_grid_s(w, h, d, len, s, edge);
  printf("E %d\n", lines);
  for (int i = 0; i < lines; i++) printf("%d %d\n", edge[i][0], edge[i][1]);
  if (!grid) ODP_Conv_edge2adjacency_general_s(nodes, lines, d, edge, s, adj);
  else ODP_Conv_edge2adjacency_grid_s(w, h, lines, d, edge, s, (int (*)[d])adj);
  printf("H %llu\n", fnv(adj, (long)rows * d));
  for (int m = 0; m < moves; m++) {
    ODP_Restore r;
    if (!grid) ODP_Mutate_adjacency_general_s(nodes, d, NULL, s, &r, (int (*)[d])adj);
    else if (s == 1 && !use_s) ODP_Mutate_adjacency_grid(w, h, d, NULL, len, &r, (int (*)[d])adj);
    else ODP_Mutate_adjacency_grid_s(w, h, d, NULL, len, s, &r, (int (*)[d])adj);
    printf("M %d %d %d %d %d %d %d %d\n", r.u[0], r.u[1], r.v[0], r.v[1], r.u_d[0], r.u_d[1], r.v_d[0], r.v_d[1]);
    if (m % 3 == 2) {
      if (!grid) ODP_Restore_adjacency_general(r, adj);
      else ODP_Restore_adjacency_grid(r, adj);
      printf("R\n");
    }
    printf("H %llu\n", fnv(adj, (long)rows * d));
  }
  printf("A\n");
  for (int i = 0; i < rows; i++) {
    for (int j = 0; j < d; j++) printf(j ? " %d" : "%d", adj[i * d + j]);
    printf("\n");
  }
  free(edge);
  free(adj);
  return 0;
}
