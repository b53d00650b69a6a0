/*
 * ref_kind.c -- runs the REFERENCE's algorithm-kind rule (utils.c:1206-1266,
 * compiled from the reference sources by oracle/Makefile into oracle/_ref/) on
 * a list of shapes.  TEST INFRASTRUCTURE ONLY: it pins the library's own
 * restatement of the rule (ODPX_Get_kind / ODPX_Get_mem_usage).
 *
 * stdin: one shape per line, "nodes degree has_num_degrees symmetries grid_s".
 * stdout: per shape "<kind> <MATRIX estimate as %a> <MATRIX_SAVING estimate as %a>",
 * the GPU branch (is_cpu = false) with one process, as the GPU flavour calls it
 * (aspl_cuda.cu:137-141).  The kind follows ODP_ASPL in the environment.
 */
#include <stdbool.h>
#include <stdio.h>

int ODP_Get_kind(const int nodes, const int degree, const int *num_degrees, const int symmetries,
                 const int procs, const bool is_cpu, const bool enable_grid_s);
double ODP_Get_mem_usage(const int kind, const int nodes, const int degree, const int symmetries,
                         const int *num_degrees, const int procs, const bool is_cpu, const bool enable_grid_s);

int main(void)
{
  int nodes, degree, has_nd, s, grid_s;
  static int dummy[1];
  while (scanf("%d %d %d %d %d", &nodes, &degree, &has_nd, &s, &grid_s) == 5) {
    const int *nd = has_nd ? dummy : NULL;
    int kind = ODP_Get_kind(nodes, degree, nd, s, 1, false, grid_s != 0);
    double m1 = ODP_Get_mem_usage(1, nodes, degree, s, nd, 1, false, grid_s != 0);
    double m2 = ODP_Get_mem_usage(2, nodes, degree, s, nd, 1, false, grid_s != 0);
    printf("%d %a %a\n", kind, m1, m2);
  }
  return 0;
}
