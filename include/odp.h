/*
 * odp.h -- C ABI of libodp_hip, a drop-in for the reference library's
 * include/odp.h.  A program written against the reference compiles against
 * this header unchanged and links with -lodp_hip instead of libodp_cuda.a
 * (sample/Makefile:25-30 picks the flavour at link time the same way).
 *
 * Every symbol keeps the reference's name, argument order, argument meaning
 * and error behaviour (invalid parameters print to stderr and exit(1),
 * common.h:32).  The comment on each declaration cites the reference
 * declaration it replaces and the definition whose behaviour it follows.
 */
#ifndef ODP_HIP_ODP_H
#define ODP_HIP_ODP_H
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Undo record of one 2-edge swap (include/odp.h:4-6). */
typedef struct {
  int u[2], v[2], u_d[2], v_d[2];
} ODP_Restore;

/* ---- ASPL / diameter lifecycle: the hot path ----------------------------
 * One evaluator per process, as in the reference (aspl_cuda.cu:2-8).
 * The results are the reference's bit-matrix method's (aspl.c:35-119;
 * aspl_cuda.cu:53-126), bit-identical to the reference's serial library.
 * Init may be called again after Finalize (the reference double-frees then,
 * aspl.c:295-296; this library resets all state at Finalize).
 * Environment, read at Init: ODP_ASPL=MATRIX|MATRIX_SAVING|BFS (utils.c:1238-1266;
 * BFS, which the reference runs on the CPU only (aspl.c:166-221) and whose GPU
 * flavour refuses (utils.c:1256-1259), runs one breadth-first search per
 * source on the GPU here, with the same results), ODP_PROFILE=1
 * (utils.c:142-150, printed at Finalize as utils.c:1268-1294 does). */

/* include/odp.h:8 ; aspl_cuda.cu:159-162 */
void ODP_Init_aspl_general(int nodes, int degree, int *num_degrees);
/* include/odp.h:9 ; aspl_cuda.cu:164-179 (num_degrees has nodes/symmetries entries) */
void ODP_Init_aspl_general_s(int nodes, int degree, int *num_degrees, int symmetries);
/* include/odp.h:10 ; aspl_cuda.cu:181-186 */
void ODP_Init_aspl_grid(int width, int height, int degree, int *num_degrees);
/* include/odp.h:11 ; aspl_cuda.cu:188-224 (num_degrees has nodes/symmetries entries) */
void ODP_Init_aspl_grid_s(int width, int height, int degree, int *num_degrees, int symmetries);
/* include/odp.h:12 ; aspl_cuda.cu:242-263.  adjacency = int[nodes/symmetries][degree];
 * disconnected graphs give INT_MAX / LONG_MAX / DBL_MAX (aspl_cuda.cu:256-260). */
void ODP_Set_aspl(void *adjacency, int *diameter, long *sum, double *ASPL);
/* include/odp.h:13 ; aspl_cuda.cu:226-240 */
void ODP_Finalize_aspl(void);

/* ---- graph utilities (host code; the reference exports them from utils.o,
 *      Makefile:38, so callers of the GPU flavour link them too) --------- */

/* include/odp.h:24 ; utils.c:1432-1453 */
void ODP_Set_lbounds_general(int nodes, int degree, int *low_diameter, double *low_ASPL);
/* include/odp.h:25 ; utils.c:1456-1504 */
void ODP_Set_lbounds_grid(int m, int n, int degree, int length, int *low_diameter, double *low_ASPL);
/* include/odp.h:26 ; utils.c:1553-1563 */
void ODP_Set_degrees(int nodes, int lines, int (*edge)[2], int *num_degrees);
/* include/odp.h:27 ; utils.c:1370-1384 */
int  ODP_Get_lines(char *fname);
/* include/odp.h:28 ; utils.c:1359-1368 */
int  ODP_Get_nodes(int lines, int (*edge)[2]);
/* include/odp.h:29 ; utils.c:1341-1357 */
int  ODP_Get_degree(int nodes, int lines, int (*edge)[2]);
/* include/odp.h:30 ; utils.c:1319-1326 */
int  ODP_Get_length(int lines, int height, int (*edge)[2]);
/* include/odp.h:31 ; utils.c:1328-1339 */
bool ODP_Check_general(char *fname);
/* include/odp.h:32 ; utils.c:1307-1317 */
bool ODP_Check_multiple_edges(int lines, int (*edge)[2]);
/* include/odp.h:33 ; utils.c:1296-1305 */
bool ODP_Check_loop(int lines, int (*edge)[2]);
/* include/odp.h:34 ; utils.c:1386-1400 */
void ODP_Read_edge_general(char *fname, int (*edge)[2]);
/* include/odp.h:35 ; utils.c:1402-1429 */
void ODP_Read_edge_grid(char *fname, int *w, int *h, int (*edge)[2]);
/* include/odp.h:36 ; utils.c:1153-1164 */
void ODP_Write_edge_general(int lines, int (*edge)[2], char *fname);
/* include/odp.h:37 ; utils.c:1166-1179 */
void ODP_Write_edge_grid(int lines, int height, int (*edge)[2], char *fname);
/* include/odp.h:38 ; utils.c:1541-1545 */
void ODP_Conv_edge2adjacency_general(int nodes, int lines, int degree, int (*edge)[2], void *adjacency);
/* include/odp.h:39 ; utils.c:1547-1551 */
void ODP_Conv_edge2adjacency_grid(int width, int height, int lines, int degree, int (*edge)[2], void *adjacency);
/* include/odp.h:40 ; utils.c:1012-1034 */
void ODP_Conv_edge2adjacency_general_s(int nodes, int lines, int degree, int (*edge)[2], int symmetries, void *adjacency);
/* include/odp.h:41 ; utils.c:1116-1151 (adjacency is int[nodes/symmetries][degree]) */
void ODP_Conv_edge2adjacency_grid_s(int width, int height, int lines, int degree, int (*edge)[2], int symmetries, void *adjacency);
/* include/odp.h:42 ; utils.c:1506-1533 */
void ODP_Conv_adjacency2edge_general(int nodes, int degree, int *num_degrees, void *adjacency, int (*edge)[2]);
/* include/odp.h:43 ; utils.c:1535-1539 */
void ODP_Conv_adjacency2edge_grid(int width, int height, int degree, int *num_degrees, void *adjacency, int (*edge)[2]);
/* include/odp.h:44 ; utils.c:991-1010 */
void ODP_Conv_adjacency2edge_general_s(int nodes, int degree, int *num_degrees, void *adjacency, int symmetries, int (*edge)[2]);
/* include/odp.h:45 ; utils.c:152-193 */
void ODP_Conv_adjacency2edge_grid_s(int width, int height, int degree, int *num_degrees, void *adjacency, int symmetries, int (*edge)[2]);
/* include/odp.h:51 ; utils.c:25-36 */
void ODP_Print_adjacency(int nodes, int degree, int *num_degrees, void *adjacency);
/* include/odp.h:52 ; utils.c:38-42 */
void ODP_Print_edge_general(int lines, int (*edge)[2]);
/* include/odp.h:53 ; utils.c:44-50 */
void ODP_Print_edge_grid(int lines, int height, int (*edge)[2]);

/* ---- random start graphs and 2-opt moves (the step before the evaluator) ----
 * Every random decision is a libc rand() draw taken as the reference takes it,
 * so one seed gives the same graphs and the same moves as the reference
 * library.  The adjacency arguments are the reference's int[rows][degree]
 * tables (rows = nodes/symmetries), passed as plain pointers. */

/* include/odp.h:46 ; utils.c:1036-1039 */
void ODP_Srand(unsigned int seed);
/* include/odp.h:47 ; utils.c:1065-1084 (edge holds nodes*degree/2 lines) */
void ODP_Generate_random_general(int nodes, int degree, int (*edge)[2]);
/* include/odp.h:48 ; utils.c:1086-1114 */
void ODP_Generate_random_general_s(int nodes, int degree, int symmetries, int (*edge)[2]);
/* include/odp.h:49 ; utils.c:1577-1625 */
void ODP_Generate_random_grid(int width, int height, int degree, int length, int (*edge)[2]);
/* include/odp.h:50 ; utils.c:1627-1750 */
void ODP_Generate_random_grid_s(int width, int height, int degree, int length, int symmetries, int (*edge)[2]);
/* include/odp.h:54 ; utils.c:784-788 */
void ODP_Mutate_adjacency_general(int nodes, int degree, int *num_degrees, ODP_Restore *r, void *adjacency);
/* include/odp.h:55 ; utils.c:752-764 */
void ODP_Mutate_adjacency_general_s(int nodes, int degree, int *num_degrees, int symmetries, ODP_Restore *r,
                                    void *adjacency);
/* include/odp.h:56 ; utils.c:1565-1574 */
void ODP_Mutate_adjacency_grid(int width, int height, int degree, int *num_degrees, int length, ODP_Restore *r,
                               void *adjacency);
/* include/odp.h:57 ; utils.c:973-989 */
void ODP_Mutate_adjacency_grid_s(int width, int height, int degree, int *num_degrees, int length, int symmetries,
                                 ODP_Restore *r, void *adjacency);
/* include/odp.h:58 ; utils.c:500-507 (the table of the last general mutation) */
void ODP_Restore_adjacency_general(ODP_Restore r, void *adjacency);
/* include/odp.h:59 ; utils.c:461-467 (the table of the last ODP_Mutate_adjacency_grid_s;
 * like the reference, nothing is stored when no such mutation came first) */
void ODP_Restore_adjacency_grid(ODP_Restore r, void *adjacency);

#ifdef __cplusplus
}
#endif
#endif
