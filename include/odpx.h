/*
 * odpx.h -- extensions of libodp_hip beyond the reference's API.  None of these
 * symbols exists in the reference; they never change the behaviour of the
 * ODP_* entry points declared in odp.h.
 *
 *  - seeded synthetic graphs and the 2-opt swap (host code);
 *  - plan control, autotuning and kernel timing of the ODP_* evaluator;
 *  - device-resident evaluation (the adjacency already in HBM);
 *  - batched evaluation of many graphs of uneven size in one persistent launch;
 *  - simulated annealing with the adjacency kept on the device;
 *  - a serving engine with continuous batching of evaluation requests (serve.hpp).
 * Every evaluation returns the same integers as ODP_Set_aspl would.
 */
#ifndef ODP_HIP_ODPX_H
#define ODP_HIP_ODPX_H
#include <stdint.h>
#include "odp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded graphs (odp_gen.c) ---------------------------------------- */
void     ODPX_Rng_seed(uint64_t seed, uint64_t state[4]);
uint64_t ODPX_Rng_next(uint64_t state[4]);
/* random simple connected degree-regular graph -> adjacency int[nodes][degree]; 0 on success */
int ODPX_Generate_regular(int nodes, int degree, uint64_t seed, int *adjacency);
/* rotationally symmetric grid graph (s in {1,2,4}) with edge length <= length -> full table int[w*h][degree] */
int ODPX_Generate_grid(int width, int height, int degree, int length, int symmetries,
                       uint64_t seed, int iters_per_edge, int *adjacency);
/* first-quadrant / left-half rows of a symmetric grid table, as ODP_Init_aspl_grid_s takes them */
int ODPX_Grid_base_table(int width, int height, int degree, int symmetries, const int *full, int *base);
/* circulant-symmetric general graph -> base table int[nodes/symmetries][degree] */
int ODPX_Generate_general_s(int nodes, int degree, int symmetries, uint64_t seed, int *base);
/* replace edges a-b, c-d by a-c, b-d in a full table */
int ODPX_Swap_edges(int nodes, int degree, int *adjacency, int a, int b, int c, int d);
/* one random valid 2-opt move; swap[4] receives {a, b, c, d} */
int ODPX_Two_opt(int nodes, int degree, int *adjacency, uint64_t rng[4], int swap[4]);

/* ---- evaluator plans -------------------------------------------------- */
#define ODPX_PLAN_AUTO     0
#define ODPX_PLAN_RESIDENT 1   /* LDS-resident blocks, one persistent launch      */
#define ODPX_PLAN_LEVEL    2   /* one launch per BFS level over HBM bit buffers   */

typedef struct {
  int plan;       /* ODPX_PLAN_RESIDENT or ODPX_PLAN_LEVEL                          */
  int wb;         /* resident: 32-bit words per row of a block (1, 2, 4 or 8); the
                     block's n x wb words must fit LDS                                 */
  int threads;    /* resident: threads per workgroup (256, 512 or 1024); at most 40
                     (wb 1), 20 (wb 2), 10 (wb 4) or 2 (wb 8) rows a thread          */
  int kw;         /* level: 64-bit words per row of a column block; 0 = whole row    */
  int upw;        /* level: row segments per wave per launch                         */
  int seg_major;  /* level: 1 = all rows of one column segment before the next        */
  int xcd;        /* level: 1 = each XCD takes one contiguous range of workgroups     */
  int group;      /* level: column blocks resident at once; 0 = all (MATRIX)          */
  int direct;     /* level: levels 0..K-1 by direct walk enumeration in one launch:
                     0 = off, -1 = the library picks K, K >= 2 = that K (at most 6)     */
  int batch;      /* level: neighbour rows gathered per batch: 4 (also 0), 8 or 16     */
  int tile;       /* level, grid graphs (kw = 2): 1 = temporal tiles, several levels per
                     launch in LDS when the entries are short enough; 0 = off           */
  int parts;      /* level, 16-word blocks: 0 = one gather per row segment (k_level);
                     2 or 4 = the XCD-phased gather over that many row parts of
                     the input slab (k_pgather)                                         */
} ODPX_Plan;

/* Plan of the current evaluator (after Init).  Set: 0 on success, -1 if invalid.
 * Without a call, Init picks one (env ODP_HIP_PLAN=resident|level overrides). */
int  ODPX_Get_plan(ODPX_Plan *plan);
int  ODPX_Set_plan(const ODPX_Plan *plan);
/* Times every candidate plan on `adjacency` (same layout as ODP_Set_aspl),
 * checks that all of them return identical results (exit(1) otherwise), and
 * keeps the fastest.  Returns the number of candidates tried, or -1 under the
 * BFS kind (which runs no plan). */
int  ODPX_Autotune(void *adjacency, int repeats);

/* ---- device-resident evaluation ---------------------------------------- */
/* Uploads adjacency to the evaluator's device copy without evaluating. */
void ODPX_Upload_adjacency(void *adjacency);
/* Evaluates the adjacency already on the device (after ODP_Set_aspl or
 * ODPX_Upload_adjacency, and any ODPX_Swap_device since). */
void ODPX_Set_aspl_resident(int *diameter, long *sum, double *ASPL);
/* Applies edge swap a-b, c-d -> a-c, b-d to the device copy (s == 1 only),
 * all or nothing: 0 on success; -1, with the graph unchanged, if an id is out of
 * range, the four ids are not distinct, or edge a-b or c-d is missing.  After
 * an upload the library keeps a host copy of the device adjacency, checks the
 * swap on it and returns without waiting for the device; otherwise (after
 * device-only changes such as ODPX_Anneal) it reads the device's verdict. */
int  ODPX_Swap_device(int a, int b, int c, int d);

/* ---- one process's share (the reference's MPI+CUDA division) ------------- */
/* Evaluates, on the active Init, only the source columns that rank `rank` of
 * `nranks` processes owns in the reference's MPI+CUDA flavour, on the adjacency
 * already on the device (after ODP_Set_aspl or ODPX_Upload_adjacency): MATRIX,
 * one block of ceil(W / nranks) 64-bit words per rank (aspl_mpi_cuda.cu:47-54,
 * W = ceil((nodes/symmetries)/64)); MATRIX_SAVING, the 64-word blocks rank,
 * rank + nranks, ... (:92-94).  Symmetric grids number the source columns as
 * the reference does (utils.c:2103-2124).  Out: excess, the rank's share of
 * the ordered-pair distance sum beyond the n(n-1) base (the reference's *sum
 * before its Allreduce, :83); diameter, the rank's largest terminal level + 1
 * (1 with no column, nodes + 1 with a block that never completes);
 * disconnected, 1 in that case.  The combine of :82-87 gives ODP_Set_aspl's
 * result: diameter = max over ranks, sum = sum over ranks + n(n-1), ASPL =
 * sum / ((n-1) n), sum /= 2, and diameter > nodes gives the sentinels.  One
 * GPU runs any rank.  Returns 0, or -1 for an invalid rank or null output. */
int  ODPX_Set_aspl_shard(int rank, int nranks, long long *excess, int *diameter, int *disconnected);

/* ---- allocation-free evaluation ----------------------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc).  ODP_Set_aspl's
 * result for a general graph (s = 1, entries -1 are absent edges) of at most
 * 16384 vertices and table width 1..32, without an Init: no hipMalloc, hipFree
 * or host heap allocation.  Every device buffer (the adjacency copy, the
 * 16-bit neighbour table, the block list, per-workgroup totals, counters and
 * the result) lives in one 3.1 MiB file-scope device array that each GPU holds
 * until the process ends; a call clears what must start at zero, and calls
 * that overlap on one GPU race.  adjacency int[nodes][degree] in host memory.
 * Returns 0, or -1 (with a message, the device untouched) for sizes past the
 * maximum; an id outside [-1, nodes) exits as ODP_Set_aspl does. */
int  ODPX_Set_aspl_noalloc(int nodes, int degree, const int *adjacency, int *diameter, long *sum, double *ASPL);

/* ---- algorithm kind (host only, no GPU) ---------------------------------- */
/* The kind ODP_Init_aspl_* picks: ODP_ASPL if set (MATRIX, MATRIX_SAVING, BFS;
 * unknown values exit as at Init), else MATRIX when the GPU memory estimate
 * of MATRIX is at most 2048 MB, MATRIX_SAVING otherwise (utils.c:1238-1266 with
 * is_cpu = false, procs = 1).  Returns 1 (MATRIX), 2 (MATRIX_SAVING) or 3 (BFS). */
int    ODPX_Get_kind(int nodes, int degree, int has_num_degrees, int symmetries, int grid_s);
/* The memory estimate in MB of that rule and of the ODP_PROFILE report
 * (utils.c:1206-1236 with is_cpu = false, procs = 1). */
double ODPX_Get_mem_usage(int kind, int nodes, int degree, int has_num_degrees, int symmetries, int grid_s);

/* ---- kernel timing ------------------------------------------------------ */
typedef struct {
  long long evals;            /* evaluations since reset                           */
  double    eval_ms;          /* wall time of those evaluations (host clock)       */
  long long level_launches;   /* gather-level launches that did work               */
  double    level_ms;         /* their summed duration, HIP events on the launch stream */
  double    level_bytes;      /* their algorithmic bytes: rows*((d+2)*k*8 + 4d)     */
  double    level_min_bytes;  /* their compulsory bytes: every row of a working block
                                 read once and (unless count-only) written once,
                                 rows*k*8*(1 + written), plus the ids once per launch */
  long long idle_launches;    /* gather-level launches after every block finished  */
  double    idle_ms;
  long long resident_launches;
  double    resident_ms;
  double    resident_bytes;   /* same formula over the levels the blocks ran       */
  long long other_launches;   /* prepare + level-0 kernels                         */
  double    other_ms;
} ODPX_Stats;
/* on: 0 off; 1 every level whole (no count-only or sparse last level); 2 the
 * launch sequence of untimed evaluations (count-only, listing and sparse last
 * levels as they come).  Timed evaluations launch eagerly (no hipGraph replay)
 * with HIP events around every launch on the library's stream. */
void ODPX_Stats_enable(int on);
void ODPX_Stats_reset(void);
void ODPX_Stats_get(ODPX_Stats *out);

/* ---- batched evaluation ------------------------------------------------- */
/* Uploads ngraphs general (s = 1) graphs, packed as for ODPX_Batch_eval, and
 * builds their neighbour tables; they stay on the device until the next load
 * or ODPX_Batch_release.  Returns 0. */
int ODPX_Batch_load(int ngraphs, const int *nodes, const int *degrees, const long long *offsets,
                    const int *packed_adjacency);
/* Evaluates the loaded graphs: every graph that fits the LDS-resident kernel in
 * one persistent launch, the others with the single-graph path.  Results
 * indexed like the loaded graphs.  Returns 0. */
int ODPX_Batch_run(int *diameter, long *sum, double *ASPL);
/* Evaluates the loaded graphs one at a time with the single-graph path
 * (plan ODPX_PLAN_LEVEL or ODPX_PLAN_AUTO): the comparison point. */
int ODPX_Loop_run(int plan, int *diameter, long *sum, double *ASPL);
/* ODPX_Batch_load + ODPX_Batch_run.  Graph g has nodes[g] vertices, table
 * width degrees[g], and its adjacency int[nodes[g]][degrees[g]] starts at
 * packed_adjacency + offsets[g]; -1 entries are absent edges.  Results as
 * ODP_Set_aspl.  Returns 0. */
int ODPX_Batch_eval(int ngraphs, const int *nodes, const int *degrees, const long long *offsets,
                    const int *packed_adjacency, int *diameter, long *sum, double *ASPL);
/* ODPX_Batch_load + ODPX_Loop_run. */
int ODPX_Loop_eval(int ngraphs, const int *nodes, const int *degrees, const long long *offsets,
                   const int *packed_adjacency, int plan, int *diameter, long *sum, double *ASPL);
/* Launch policy of ODPX_Batch_run: wbcap[5] = the widest LDS block (32-bit
 * words per row: 1, 2, 4 or 8) for graphs with n <= 2048, 4096, 8192, 16384
 * and larger; per_cu = resident workgroups per CU (1 or 2).  Set returns 0, or
 * -1 for an invalid policy; it rebuilds the loaded batch's task table. */
void ODPX_Batch_get_policy(int *wbcap, int *per_cu);
int  ODPX_Batch_set_policy(const int *wbcap, int per_cu);
/* Autotunes the policy on the loaded batch: times candidates (workgroups per
 * CU, then the block width per size class, coordinate descent), checks that
 * all return identical results (exit(1) otherwise) and keeps the fastest.
 * Returns the number of candidates timed. */
int  ODPX_Batch_autotune(int repeats);
/* Frees the device buffers the batched and looped paths keep between calls. */
void ODPX_Batch_release(void);

/* ---- simulated annealing on the device copy ----------------------------- */
typedef struct {
  long long steps;            /* moves evaluated                                      */
  long long accepted;         /* moves the acceptance test kept                       */
  int    best_diameter;       /* the best graph seen (returned in adjacency)          */
  long   best_sum;
  double best_ASPL;
  int    diameter;            /* the chain's current graph at the end                 */
  double ASPL;
  int    optimum;             /* 1: the lower bounds of ODP_Set_lbounds_general reached */
  double seconds;
} ODPX_AnnealResult;
/* One 2-opt move of the reference's kind (the move of ODP_Mutate_adjacency_*,
 * utils.c:571-990, with its rejections and symmetric images) on a base table,
 * drawn from a seeded generator instead of libc rand(): family 0 general
 * (nodes_or_width = nodes), 1 grid (width, height; edges at most `length` long).
 * The undo record goes to r, for ODP_Restore_adjacency_general / _grid.
 * Returns 0, or -1 for invalid parameters. */
int ODPX_Mutate(int family, int nodes_or_width, int height, int degree, int length, int symmetries,
                void *adjacency, uint64_t rng[4], ODP_Restore *r);
/* The annealing loop of the reference's sample/SA/general.c:149-181 for a general
 * graph (an active ODP_Init_aspl_general(nodes, degree, ...)): ncalcs times,
 * one random 2-opt move (ODPX_Two_opt, generator seeded by seed) applied to the
 * device copy, an exact evaluation (as ODP_Set_aspl), the best graph kept, and the
 * acceptance test of sample/SA/common.c:15-43 (temperature from max_temp to
 * min_temp geometrically; hill_climbing accepts no worse ASPL) -- a rejected move
 * is undone on the device.  adjacency int[nodes][degree]: in, the start graph;
 * out, the best graph.  Stops early at the lower bounds.  0, or -1 if invalid.
 * Under the resident plan, a regular table whose block and neighbour table fit
 * one workgroup's LDS runs many steps per cooperative launch with no host round
 * trip per step (k_anneal: every workgroup draws the same move from a copy of
 * the generator); ODP_HIP_ANNEAL=host keeps one launch per step.
 * Symmetric general (general_s) and grid Inits anneal the base table the Init
 * takes with ODPX_Mutate's moves (grid edges kept within the start graph's
 * longest Manhattan length), the acceptance test accept_s (sample/SA/common.c:
 * 15-37, the difference divided by symmetries) and the reference's restores. */
int ODPX_Anneal(void *adjacency, long long ncalcs, double max_temp, double min_temp, int hill_climbing,
                int ASPL_priority, uint64_t seed, ODPX_AnnealResult *result);

/* ---- serving: continuous batching of 2-opt evaluation requests --------- */
/* A chain is one graph (general, s = 1) kept in HBM; a request applies a list
 * of 2-opt swaps (a-b, c-d -> a-c, b-d, each as {a, b, c, d}) to its chain and
 * evaluates the result exactly.  Requests of one chain run in submission order;
 * requests of different chains are batched and overlapped on the GPU as
 * slots free up.  Results equal ODP_Set_aspl on the same graph. */
typedef struct {
  long long id;               /* request id returned by ODPX_Serve_submit            */
  int    chain;
  int    status;              /* 0, or -1 if a swap named an edge the graph lacks    */
  int    diameter;
  long   sum;
  double ASPL;
  double t_submit, t_start, t_done;   /* seconds on the host's monotonic clock     */
} ODPX_Result;
/* small_slots: concurrent LDS-resident batch launches (each its own stream);
 * large_slots: concurrent evaluations of graphs too large for LDS, each holding
 * its own bit buffers and stream; hbm_budget_gb: bound on chain tables plus
 * those buffers.  0 on success.  The server is advanced by its caller: submit,
 * poll, wait and drain collect finished work and admit waiting requests, so
 * every HIP call is made by the calling thread (calls are serialised). */
int  ODPX_Serve_start(int small_slots, int large_slots, double hbm_budget_gb);
/* adjacency int[nodes][degree]; returns the chain id, or -1 (over the HBM budget,
 * invalid graph, or no server) */
int  ODPX_Serve_add_chain(int nodes, int degree, const int *adjacency);
/* returns the request id (>= 1) or -1 */
long long ODPX_Serve_submit(int chain, int nswaps, const int *swaps);
/* up to max completed results, in completion order; returns how many */
int  ODPX_Serve_poll(ODPX_Result *out, int max);
/* blocks until request id completes; 0, or -1 if the id is unknown */
int  ODPX_Serve_wait(long long id, ODPX_Result *out);
/* blocks until every submitted request has completed */
void ODPX_Serve_drain(void);
/* drains, frees the slots and the chains */
void ODPX_Serve_stop(void);

#ifdef __cplusplus
}
#endif
#endif
