// kernels.hpp -- gfx950 (CDNA4, wave64) kernels of the bit-matrix ASPL/diameter
// method.  Included once, by aspl_hip.hip.
//
// The method is the reference's (aspl.c:35-119, aspl_cuda.cu:53-126,
// utils_cuda.cu:97-211): bit c of row v says "source c reaches v"; each level
// computes next[v] = cur[v] | OR_{u in N(v)} cur[u] on 64-bit words and
// popcounts next.  Two kernel families implement it:
//
//  * level kernels (k_level0, k_level): one launch per BFS level for every
//    column block ("task") of one or more graphs.  A wave owns a segment of
//    64..128 words of a row (16 B per lane, so every neighbour gather is one
//    coalesced 512 B..1 KiB stream).  The popcount is fused and reduced per
//    wave into one 64-bit atomic; termination (count == n*l) is detected on
//    the device by the next level's workgroups, so the host needs no sync per
//    level (the reference syncs every level, aspl_cuda.cu:67).  Counts live in
//    a ring of `ring` slots per task; level L zeroes slot L+1 before level L+1
//    adds into it, and the host reads the ring between batches of levels.
//  * resident kernel (k_resident): for graphs small enough that a block of
//    32*wb sources of BOTH bit buffers fits in the 160 KiB LDS of a CU,
//    one workgroup runs every level of one block in LDS.  Workgroups pull
//    blocks from a work queue (one atomic counter), so many graphs of uneven
//    size are evaluated in one persistent launch.
//
// Neighbour ids reaching these kernels come from the row-neighbour table
// built by k_prepare: every entry is a valid row index (short rows are padded
// with the row's own index, which ORs in nothing new), so no gather can leave
// the bit buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;
typedef unsigned int u32;

struct __attribute__((aligned(16))) u64x2 { u64 x, y; };
typedef u64 u64v2 __attribute__((ext_vector_type(2)));

#define ODP_NXCD 8          // XCDs of an MI355X; workgroups are dealt round-robin
#define LEVEL_WG 256        // threads per level-kernel workgroup (4 waves)
#define LV_STORE 1          // k_level mode bits: write the next rows
#define LV_COUNT 2          //   ... and add the level's counts (both: a normal level)
#define LV_LIST 4           //   ... and list the chunks left unsaturated (k_level's LS variant)
#define LV_SPARSE 8         // the count-only level as k_sparse over that list

// chunks (task, row, 16-byte word pair) a level left unsaturated; only those
// can gain bits at the next level.  The list is LIST_SUBS sub-lists with a
// counter each (128 bytes apart), a workgroup appending to sub-list
// blockIdx.x % LIST_SUBS: one counter for the whole GPU serialised the appends
// (a listing level of n = 49152, d = 4 took 3.4 ms against 0.2 ms unlisted).
#define LIST_SUBS 64
#define LIST_PAD 32
struct LvList {
  u64* e;              // sub-list s at e + s * cap; entries: task << 48 | row << 24 | word
  u32* n;              // sub-list s's count at n[s * LIST_PAD] (may exceed cap: then no list is used)
  u32 cap;             // entries per sub-list
};

// ---------------------------------------------------------------------------
// shared helpers

// appends the wave's lanes with `unsat` set to the workgroup's sub-list: one
// atomic per wave.  Called by whole waves.
__device__ __forceinline__ void list_append(const LvList& lst, bool unsat, u64 val)
{
  const u64 m = __ballot(unsat);
  if (!m) return;
  const u32 sub = blockIdx.x % LIST_SUBS, lane = threadIdx.x & 63;
  u32 b = 0;
  if (lane == 0) b = atomicAdd(lst.n + sub * LIST_PAD, (u32)__popcll(m));
  b = __builtin_amdgcn_readfirstlane(b);
  if (unsat) {
    const u32 e = b + (u32)__popcll(m & ((1ULL << lane) - 1));
    if (e < lst.cap) lst.e[(u64)sub * lst.cap + e] = val;
  }
}

__device__ __forceinline__ u64 wave_sum(u64 v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// valid-bit mask of word w of a block holding l sources
__device__ __forceinline__ u64 full_word(int l, int w)
{
  int v = l - 64 * w;
  return v >= 64 ? ~0ULL : (v <= 0 ? 0ULL : ((1ULL << v) - 1ULL));
}

// blockIdx -> logical index so that each XCD gets one contiguous range
__device__ __forceinline__ int remap_xcd(int b, int nb)
{
  int q = nb / ODP_NXCD, r = nb % ODP_NXCD, x = b % ODP_NXCD, i = b / ODP_NXCD;
  return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------------------
// k_prepare: caller adjacency -> row-neighbour table.
//
//  s == 1              : row i, neighbour adj[i][j]                      (utils_cuda.cu:101-113)
//  general, s > 1      : row i = p*b+m, neighbour adj[m][j] + p*b (mod n) (utils_cuda.cu:133-151)
//  grid, s in {2, 4}   : vertex i lives in row itable[i]; neighbour is the
//                        rotated first-quadrant entry (utils_cuda.cu:3-59, 117-132)
// Entries j >= num_degrees[i] become the row itself.  Ids outside [0, n) set
// *err and are replaced by the row itself (the reference would read out of
// bounds; the host turns *err into the reference's print-and-exit).
// The first threads also reset the per-evaluation counters.

struct PrepArgs {
  const int* adj;        // [(n/s)][degree]
  const int* numdeg;     // [n] expanded, or null
  const int* itable;     // [n] or null
  int* nbr;              // [n][dpad] 32-bit ids, or null
  unsigned short* nbr16; // [n][dpad] 16-bit ids (n <= 65536), or null (both may be written)
  int* err;
  int n, s, degree, dpad, width, height, grid_s;
  u64* zero64; int nzero64;   // counters reset to 0 at the start of an evaluation
  u32* zero32; int nzero32;
  int* adj_copy; int64_t adj_elems;   // adj is in host memory: also copy it here (or null)
};

__device__ __forceinline__ int grid_rotate(int v, int width, int height, int s, int angle)
{
  int w = v / height, h = v % height;
  if (s == 2) return (width - w - 1) * height + (height - h - 1);
  if (angle == 90) return h * height + (height - w - 1);
  if (angle == 180) return (height - w - 1) * height + (height - h - 1);
  return (height - h - 1) * height + w;
}

// neighbour d of vertex v of a 2/4-fold symmetric grid; -1 when the table
// entry is not a vertex id
__device__ __forceinline__ int grid_nbr(const int* adj, int width, int height, int degree, int s, int v, int d)
{
  int n = width * height, bw = width / 2, bh = height / 2, w = v / height, h = v % height;
  int base, back = 0;
  if (s == 2) {
    if (w < bw) base = v;
    else { base = grid_rotate(v, width, height, 2, 180); back = 180; }
  } else {
    int x = v, lx;
    if (w < bw && h < bh) back = 0;
    else if (w < bw) { x = grid_rotate(v, width, height, 4, 270); back = 90; }
    else if (h >= bh) { x = grid_rotate(v, width, height, 4, 180); back = 180; }
    else { x = grid_rotate(v, width, height, 4, 90); back = 270; }
    lx = (x / height) * bh + (x % height);
    base = lx;
  }
  int y = adj[(int64_t)base * degree + d];
  if (y < 0 || y >= n) return -1;
  return back ? grid_rotate(y, width, height, s, back) : y;
}

__global__ void k_prepare(PrepArgs a)
{
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = tid; i < a.nzero64; i += stride) a.zero64[i] = 0;
  for (int64_t i = tid; i < a.nzero32; i += stride) a.zero32[i] = 0;
  if (a.adj_copy)
    for (int64_t i = tid; i < a.adj_elems; i += stride) a.adj_copy[i] = a.adj[i];
  const int64_t total = (int64_t)a.n * a.dpad;
  const int b = a.n / a.s;
  for (int64_t idx = tid; idx < total; idx += stride) {
    const int i = (int)(idx / a.dpad), j = (int)(idx - (int64_t)i * a.dpad);
    const int row = a.grid_s ? a.itable[i] : i;
    int d = a.numdeg ? a.numdeg[i] : a.degree;
    if (d < 0 || d > a.degree) { atomicOr(a.err, 2); d = d < 0 ? 0 : a.degree; }
    int out = row;
    if (j < d) {
      int v;
      if (a.grid_s) {
        v = grid_nbr(a.adj, a.width, a.height, a.degree, a.s, i, j);
      } else if (a.s == 1) {
        v = a.adj[(int64_t)i * a.degree + j];
        if (v >= a.n) v = -1;
      } else {
        const int p = i / b, m = i - p * b;
        v = a.adj[(int64_t)m * a.degree + j];
        if (v >= a.n) v = -1;
        if (v >= 0) { v += p * b; if (v >= a.n) v -= a.n; }
      }
      if (v < 0) atomicOr(a.err, 1);
      else out = a.grid_s ? a.itable[v] : v;
    }
    if (a.nbr16) a.nbr16[(int64_t)row * a.dpad + j] = (unsigned short)out;
    if (a.nbr) a.nbr[(int64_t)row * a.dpad + j] = out;
  }
}

// the row-neighbour tables a device swap patches along with the adjacency
struct SwapTables {
  int* t32;                 // 32-bit table or null
  unsigned short* t16;      // 16-bit table or null
  int tdp;                  // their row stride
};

// 2-opt swap a-b, c-d -> a-c, b-d on the device adjacency (s = 1), all or
// nothing: one thread finds the four entries first (rows a, b, c, d holding b,
// a, d, c) and writes only when all four exist; otherwise it changes nothing
// and sets bit 4 of *err.  The ids must be distinct (the host checks), so the
// four entries are four different cells.  The tables set in `tb` are patched
// entry for entry: a table row holds the adjacency's entry where it is a
// neighbour and the row itself where it is padding, so an entry equal to
// `from` is the same edge.
__global__ void k_swap(int* adj, int degree, int a, int b, int c, int d, int* err, SwapTables tb)
{
  __shared__ int spos[4];
  auto row = [&](int e) { return e == 0 ? a : e == 1 ? b : e == 2 ? c : d; };
  auto from = [&](int e) { return e == 0 ? b : e == 1 ? a : e == 2 ? d : c; };
  auto to = [&](int e) { return e == 0 ? c : e == 1 ? d : e == 2 ? a : b; };
  if (threadIdx.x < 4) spos[threadIdx.x] = INT_MAX;
  __syncthreads();
  // the four rows' entries are read at once; a row's first match wins
  for (int i = threadIdx.x; i < 4 * degree; i += blockDim.x) {
    const int e = i / degree, j = i - e * degree;
    if (adj[(int64_t)row(e) * degree + j] == from(e)) atomicMin(&spos[e], j);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int e = 0; e < 4; e++)
    if (spos[e] == INT_MAX) { atomicOr(err, 4); return; }
  for (int e = 0; e < 4; e++) {
    adj[(int64_t)row(e) * degree + spos[e]] = to(e);
    const int64_t k = (int64_t)row(e) * tb.tdp + spos[e];
    if (tb.t16 && tb.t16[k] == (unsigned short)from(e)) tb.t16[k] = (unsigned short)to(e);
    if (tb.t32 && tb.t32[k] == from(e)) tb.t32[k] = to(e);
  }
}

// Sets *asym when some row u is listed among the neighbours of row r while r
// is not among u's (entries equal to the row itself are padding).
__global__ void k_symcheck(const int* __restrict__ nbr, int n, int dpad, u32* asym)
{
  const int64_t total = (int64_t)n * dpad;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / dpad);
    const int u = nbr[idx];
    if (u == r) continue;
    bool found = false;
    for (int k = 0; k < dpad && !found; k++) found = nbr[(int64_t)u * dpad + k] == r;
    if (!found) atomicOr(asym, 1u);
  }
}

// ---------------------------------------------------------------------------
// level kernels

struct LevelTask {
  const int* nbr;      // [n][dpad]
  u64* buf0;           // [n][kwp]  level L reads buf[L&1], writes buf[(L+1)&1]
  u64* buf1;
  u64* cnt;            // per-level popcount of this task: cnt[(L % ring) * cstride]
  u32* done;           // set once the task has terminated
  int n, dpad, kwp;    // rows, neighbour-table width, words per row in the buffers
  int src0, l;         // first source (row index) of the block, sources in it
  int lpr_log2;        // log2(lanes per row): 6 -> one row per wave
  int segs;            // 2*64-word segments per row (kwp / (2 * lanes per row))
  int rgroups;         // ceil(n / rows per wave)
  int units;           // rgroups * segs wave-units
  int seg_major;       // unit order: segment-major (1) or row-major (0)
  int cstride;         // level stride of cnt and ccnt
  int pad;
  u64* ccnt;           // per-level count of changed (row, segment) units: ccnt[(L % ring) * cstride]
  u64* acnt;           // per-level count of processed (row, segment) units, same layout
  unsigned char* chg0; // changed flag of (segment s, row r) at [s * n + r]: levels of even parity
  unsigned char* chg1; //   ... and of odd parity
  unsigned char* act0; // "may change" marks of (segment, row) for levels of even parity (push mode)
  unsigned char* act1; //   ... and of odd parity
  const u32* asym;     // nonzero when the table is not symmetric (push mode needs u in N(r) <=> r in N(u))
  unsigned char* wgm0; // per workgroup of the task: some row of it is marked (push mode), even levels
  unsigned char* wgm1; //   ... odd levels
};

// workgroup (within its task) that processes segment seg of row r, for upw
// units per wave and 4 waves per workgroup
__device__ __forceinline__ int unit_wg(const LevelTask& T, int r, int seg, int upw)
{
  const int rg = r >> (6 - T.lpr_log2);
  const int u = T.seg_major ? seg * T.rgroups + rg : rg * T.segs + seg;
  return u / (4 * upw);
}

__device__ __forceinline__ int find_task(const int* __restrict__ wg_start, int ntasks, int g)
{
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (wg_start[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct UnitGeom { int r, w, seg; };

__device__ __forceinline__ UnitGeom unit_geom(const LevelTask& T, int u, int lane)
{
  int seg, rg;
  if (T.seg_major) { seg = u / T.rgroups; rg = u - seg * T.rgroups; }
  else { rg = u / T.segs; seg = u - rg * T.segs; }
  const int lpr = 1 << T.lpr_log2;
  UnitGeom g;
  g.r = (rg << (6 - T.lpr_log2)) + (lane >> T.lpr_log2);
  g.w = seg * (2 * lpr) + 2 * (lane & (lpr - 1));
  g.seg = seg;
  return g;
}

__device__ __forceinline__ void put_bit(u64& x0, u64& x1, int c, int l, int w)
{
  if ((unsigned)c < (unsigned)l) {
    const int cw = c >> 6;
    const u64 bit = 1ULL << (c & 63);
    if (cw == w) x0 |= bit;
    else if (cw == w + 1) x1 |= bit;
  }
}

// Level 0 of every task straight from the identity: next[r] holds the bits of
// r itself and of its neighbours (cur = identity needs no reads).
__global__ __launch_bounds__(LEVEL_WG) void k_level0(const LevelTask* __restrict__ tasks,
                                                     const int* __restrict__ wg_start, int ntasks,
                                                     int ring, int upw, int xcd)
{
  int g = blockIdx.x;
  if (xcd) g = remap_xcd(g, gridDim.x);
  const int t = ntasks > 1 ? find_task(wg_start, ntasks, g) : 0;
  const LevelTask T = tasks[t];
  const int wg = g - wg_start[t];
  if (wg == 0 && threadIdx.x == 0) T.cnt[(int64_t)(1 % ring) * T.cstride] = 0;   // level 1's slot
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  u64 pop = 0;
  const int base = wg * 4 * upw;
  for (int i = 0; i < upw; i++) {
    const int u = base + i * 4 + wave;
    if (u >= T.units) break;
    const UnitGeom q = unit_geom(T, u, lane);
    if (q.r < T.n) {
      u64 x0 = 0, x1 = 0;
      put_bit(x0, x1, q.r - T.src0, T.l, q.w);
      const int* nb = T.nbr + (int64_t)q.r * T.dpad;
      for (int j = 0; j < T.dpad; j += 4) {
        const int4 v = *(const int4*)(nb + j);
        put_bit(x0, x1, v.x - T.src0, T.l, q.w);
        put_bit(x0, x1, v.y - T.src0, T.l, q.w);
        put_bit(x0, x1, v.z - T.src0, T.l, q.w);
        put_bit(x0, x1, v.w - T.src0, T.l, q.w);
      }
      u64x2 o; o.x = x0; o.y = x1;
      *(u64x2*)(T.buf1 + (int64_t)q.r * T.kwp + q.w) = o;
      pop += (u64)(__popcll(x0) + __popcll(x1));
    }
  }
  pop = wave_sum(pop);
  if (lane == 0 && pop) atomicAdd(T.cnt, pop);
}

// One gather level L >= 1 of every task.  UROW: one row per wave, so the row's
// neighbour ids are wave-uniform (scalar loads).  G: neighbour rows gathered
// per batch (in flight together) before the wave checks for saturation.
//
// FR (chosen per evaluation, for graphs with many levels): every level records,
// per (row, segment), whether the segment changed and how many did.  For a
// symmetric table (push mode) a changed segment also marks itself and its
// neighbours' segments for level L+1, so the next level tests one byte per
// row; otherwise it tests the row's and its neighbours' changed flags (pull).
// When fewer than half changed at level L-1 (and L-1 was not the first
// gathered level, which wrote every row), level L runs in frontier
// mode: a segment whose row and neighbours all stayed unchanged at L-1 cannot
// change at L, and the buffer level L writes already holds its value (the
// level L-2 value, equal to L-1's), so it is skipped -- no reads, no write.
// The count of level L is the count of L-1 plus each processed segment's
// gain, which equals the popcount of the whole level.
//
// mode: LV_STORE | LV_COUNT for a normal level; LV_COUNT alone for a level
// expected to be the last one (its rows are never read again, so they are not
// written); LV_STORE alone to write the rows of such a level after all when
// it turned out not to be the last (its counts are already in place).  The
// LS variant (plain mode only) also appends every chunk it leaves unsaturated
// to lst (LV_LIST), for k_sparse to count the next level from.
template <bool UROW, int G, bool FR, bool LS>
__global__ __launch_bounds__(LEVEL_WG) void k_level(const LevelTask* __restrict__ tasks,
                                                    const int* __restrict__ wg_start, int ntasks,
                                                    int level, int ring, int upw, int xcd, int first_level,
                                                    int mode, LvList lst)
{
  int g = blockIdx.x;
  if (xcd) g = remap_xcd(g, gridDim.x);
  const int t = ntasks > 1 ? find_task(wg_start, ntasks, g) : 0;
  const LevelTask T = tasks[t];
  if (*T.done) return;
  const u64 target = (u64)T.n * (u64)T.l;
  const u64 c1 = T.cnt[(int64_t)((level - 1) % ring) * T.cstride];
  {
    const bool stop = (c1 == target) || (level >= 2 && c1 == T.cnt[(int64_t)((level - 2) % ring) * T.cstride]);
    if (stop) {
      if (threadIdx.x == 0) *T.done = 1;
      return;
    }
  }
  const int wg = g - wg_start[t];
  const bool count = (mode & LV_COUNT) != 0, store = (mode & LV_STORE) != 0;
  if (count && wg == 0 && threadIdx.x == 0) {
    T.cnt[(int64_t)((level + 1) % ring) * T.cstride] = 0;
    if (FR) {
      T.ccnt[(int64_t)((level + 1) % ring) * T.cstride] = 0;
      T.acnt[(int64_t)((level + 1) % ring) * T.cstride] = 0;
      atomicAdd(&T.cnt[(int64_t)(level % ring) * T.cstride], c1);    // the gains below add to it
    }
  }
  const bool fr = FR && level > first_level &&
                  2 * T.ccnt[(int64_t)((level - 1) % ring) * T.cstride] < (u64)T.n * (u64)T.segs;
  const unsigned char* __restrict__ chr = (level & 1) ? T.chg0 : T.chg1;   // written by level L-1
  unsigned char* __restrict__ chw = (level & 1) ? T.chg1 : T.chg0;
  const bool push = FR && *T.asym == 0;
  unsigned char* actr = (level & 1) ? T.act1 : T.act0;    // marks for this level
  unsigned char* actw = (level & 1) ? T.act0 : T.act1;    // marks for level L+1
  unsigned char* wgmw = (level & 1) ? T.wgm0 : T.wgm1;
  if (fr && push) {
    // a workgroup none of whose rows is marked has nothing to do at this level
    unsigned char* wgmr = (level & 1) ? T.wgm1 : T.wgm0;
    if (wgmr[wg] == 0) return;
    __syncthreads();                                   // every thread has read the mark
    if (store && threadIdx.x == 0) wgmr[wg] = 0;       // cleared like the row marks
  }
  const u64* __restrict__ cur = (level & 1) ? T.buf1 : T.buf0;
  u64* __restrict__ nxt = (level & 1) ? T.buf0 : T.buf1;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int lpr = 1 << T.lpr_log2;
  const u64 gmask = lpr == 64 ? ~0ULL : ((1ULL << lpr) - 1);
  const int64_t kwp = T.kwp;
  u64 pop = 0, nch = 0, nact = 0;
  const int base = wg * 4 * upw;
  for (int i = 0; i < upw; i++) {
    const int u = base + i * 4 + wave;
    if (u >= T.units) break;
    UnitGeom q = unit_geom(T, u, lane);
    if (UROW) q.r = __builtin_amdgcn_readfirstlane(q.r);
    bool unsat = false;
    if (q.r < T.n) {
      const int64_t so = (int64_t)q.seg * T.n;
      const int* nb = T.nbr + (int64_t)q.r * T.dpad;
      bool act = true;
      if (fr && push) {
        // marked by level L-1 for a change of this row or a neighbour; the pass
        // that writes the rows clears the mark (a count-only pass leaves it for
        // its store-only re-run), so the array is clean when level L+1 marks it
        act = actr[so + q.r] != 0;
        if (act && store && (lane & (lpr - 1)) == 0) actr[so + q.r] = 0;
      } else if (fr) {
        act = chr[so + q.r] != 0;
        for (int j = 0; j < T.dpad && !act; j += 4) {
          const int4 v = *(const int4*)(nb + j);
          act = (chr[so + v.x] | chr[so + v.y] | chr[so + v.z] | chr[so + v.w]) != 0;
        }
      }
      bool changed = false;
      if (act) {
        const u64 f0 = full_word(T.l, q.w), f1 = full_word(T.l, q.w + 1);
        const u64x2 x0 = *(const u64x2*)(cur + q.r * kwp + q.w);
        u64x2 x = x0;
        // Once a lane's words are saturated the remaining neighbours cannot add
        // a bit (OR is monotone): the lane stops gathering (its loads are masked
        // off, so they cost no memory requests) and the wave stops once every
        // lane has.
        bool need = !(x.x == f0 && x.y == f1);
        if (__any(need)) {
          for (int j = 0; j < T.dpad; j += G) {
            int id[G];
#pragma unroll
            for (int k = 0; k < G; k += 4) {
              int4 v;
              if (need && (k == 0 || j + k < T.dpad)) v = *(const int4*)(nb + j + k);
              else { v.x = q.r; v.y = q.r; v.z = q.r; v.w = q.r; }
              id[k] = v.x; id[k + 1] = v.y; id[k + 2] = v.z; id[k + 3] = v.w;
            }
            u64x2 gv[G];
#pragma unroll
            for (int k = 0; k < G; k++) {
              if (need) gv[k] = *(const u64x2*)(cur + (int64_t)id[k] * kwp + q.w);
              else { gv[k].x = 0; gv[k].y = 0; }
            }
#pragma unroll
            for (int k = 0; k < G; k++) { x.x |= gv[k].x; x.y |= gv[k].y; }
            need = !(x.x == f0 && x.y == f1);
            if (!__any(need)) break;
          }
        }
        if (store) *(u64x2*)(nxt + q.r * kwp + q.w) = x;
        if (LS) unsat = !(x.x == f0 && x.y == f1);
        if (FR) {
          changed = x.x != x0.x || x.y != x0.y;
          pop += (u64)(__popcll(x.x) + __popcll(x.y) - __popcll(x0.x) - __popcll(x0.y));
        } else {
          pop += (u64)(__popcll(x.x) + __popcll(x.y));
        }
      }
      if (FR && count) {
        // the segment changed if any of its lanes did
        const u64 m = __ballot(changed);
        const bool any = ((m >> (lane & ~(lpr - 1))) & gmask) != 0;
        if ((lane & (lpr - 1)) == 0) {
          if (!push) chw[so + q.r] = any ? 1 : 0;     // push mode reads marks, not flags
          nch += any ? 1 : 0;
          nact += act ? 1 : 0;
        }
        if (push && any) {
          // mark the row and its neighbours for level L+1 (the table is symmetric:
          // the rows that gather this one are its neighbours), and their
          // workgroups; the row's lanes share the neighbours
          if ((lane & (lpr - 1)) == 0) {
            actw[so + q.r] = 1;
            wgmw[unit_wg(T, q.r, q.seg, upw)] = 1;
          }
          for (int j = lane & (lpr - 1); j < T.dpad; j += lpr) {
            const int x = nb[j];
            actw[so + x] = 1;
            wgmw[unit_wg(T, x, q.seg, upw)] = 1;
          }
        }
      }
    }
    if (LS) list_append(lst, unsat, ((u64)t << 48) | ((u64)q.r << 24) | (u64)q.w);   // a slot per unsaturated chunk
  }
  if (!count) return;
  pop = wave_sum(pop);
  if (lane == 0 && pop) atomicAdd(&T.cnt[(int64_t)(level % ring) * T.cstride], pop);
  if (FR) {
    nch = wave_sum(nch);
    nact = wave_sum(nact);
    if (lane == 0 && nch) atomicAdd(&T.ccnt[(int64_t)(level % ring) * T.cstride], nch);
    if (lane == 0 && nact) atomicAdd(&T.acnt[(int64_t)(level % ring) * T.cstride], nact);
  }
}

// ---------------------------------------------------------------------------
// part-phased gather level (k_pgather): the plain level of k_level for blocks
// of 16 words (128-byte rows), arranged so that one XCD's L2 holds what its
// workgroups gather from.
//
// A block's input slab (n rows x 128 B: 8 MiB at n = 65536) does not fit the
// 4 MB L2 of an XCD, and random neighbours spread the gathers over all of it.
// So the rows of the slab are cut into P parts (rows [p n / P, (p+1) n / P)).
// Rows are taken in units of PG_UNIT consecutive rows; for every unit and part
// a list holds the unit's (row, neighbour) pairs whose neighbour lies in that
// part (the row itself counts as a neighbour; repeats are dropped), built by
// k_plist with the table (a device swap rebuilds the units of the rows it
// changes).  Workgroups are persistent and grouped
// by XCD (workgroup b runs on XCD b % 8): the workgroups of one XCD take
// consecutive units -- the same block -- and sweep the parts in order, with
// their rows' next values accumulated in LDS (ds_or) across the sweep.  While
// they are in part p, the XCD's gathers fall into one n/P-row piece of the slab
// (2 MiB at P = 4), which its L2 keeps.  The gathers need no masking (every
// pair is a 128-byte row read) and none waits for another.  Counts, stop test
// and listing are k_level's (the plain variant): every result is bit-identical.

#define PG_MAXP 4
#ifndef PG_SLOTS
#define PG_SLOTS 12          // gathers a wave keeps in flight (registers: 4 per slot)
#endif
#define PG_CH (8 * PG_SLOTS) // pairs of a chunk
#define PG_NL ((PG_CH + 63) / 64)
#define PG_UNIT 1024         // rows of a unit, 128 KiB of LDS rows: one workgroup per CU
#ifndef PG_WG
#define PG_WG 1024           // threads of the workgroup (a divisor of PG_UNIT)
#endif

struct PgArgs {
  const u32* list;       // per unit, parts in order: (row in unit) << 22 | neighbour
  const int* poff;       // per unit: P + 1 offsets into its list (relative to unit * PG_UNIT * sp)
  int n, sp, P;
  int upt;               // units per task: ceil(n / PG_UNIT)
  long long units;       // tasks x upt
};

// One workgroup of PG_UNIT threads per unit, a thread per row: the row's
// distinct neighbours and the row itself (padding entries are the row), their
// parts, a scan of the rows' run lengths per part, and the unit's pairs
// grouped by part (in any order within a part) with the part offsets.
// DMAX >= dpad + 1 bounds the ids a row holds in registers.
//
// A device swap (or a small change of an uploaded table) hands the changed
// entries to the units it rebuilds (rows, positions in the row, old and new
// ids), so that they see them whether or not the table already holds them.
#define PL_MOVE_MAX 8
struct PlMove { int row[PL_MOVE_MAX], pos[PL_MOVE_MAX], from[PL_MOVE_MAX], to[PL_MOVE_MAX]; int cnt; };

// one unit's pair lists (a workgroup of PG_UNIT threads, unit u)
template <int DMAX>
__device__ __forceinline__ void plist_unit(const int* __restrict__ nbr, int n, int dpad, int sp, int P,
                                           u32* __restrict__ list, int* __restrict__ poff, int u, const PlMove& mv)
{
  __shared__ u64 wtot[PG_UNIT / 64];
  const int t = threadIdx.x, r = u * PG_UNIT + t, lane = t & 63, wave = t >> 6;
  // the row's table entries, 16 bytes a load (dpad is a multiple of 4), every
  // thread loading (a row past n reads the last row's): a load under a lane
  // mask would make the compiler wait for each one before the next
  int id[DMAX];
  const int4* src = (const int4*)(nbr + (int64_t)(r < n ? r : n - 1) * dpad);
#pragma unroll
  for (int j4 = 0; j4 < DMAX / 4; j4++) {
    int4 v = make_int4(-1, -1, -1, -1);
    if (4 * j4 < dpad) v = src[j4];                 // uniform
    id[4 * j4] = v.x; id[4 * j4 + 1] = v.y; id[4 * j4 + 2] = v.z; id[4 * j4 + 3] = v.w;
  }
#pragma unroll
  for (int j = 0; j < DMAX; j++)
    if (r >= n || j > dpad) id[j] = -1;
    else if (j == dpad) id[j] = r;
  for (int e = 0; e < mv.cnt; e++)
    if (r == mv.row[e]) {
#pragma unroll
      for (int j = 0; j < DMAX; j++)
        if (j == mv.pos[e] && id[j] == mv.from[e]) id[j] = mv.to[e];
    }
  // part of an id: floor(id P / n), the number of cuts ceil(p n / P) (0 < p < P) at or below it
  int cut[PG_MAXP - 1];
#pragma unroll
  for (int p = 1; p < PG_MAXP; p++) cut[p - 1] = p < P ? (int)(((int64_t)p * n + P - 1) / P) : INT_MAX;
  auto part = [&](int v) {
    int q = 0;
#pragma unroll
    for (int p = 0; p < PG_MAXP - 1; p++) q += v >= cut[p] ? 1 : 0;
    return q;
  };
  // drop repeats (an id equal to an earlier one), count the rest per part
#pragma unroll
  for (int j = 1; j < DMAX; j++) {
    bool dup = false;
#pragma unroll
    for (int k = 0; k < j; k++) dup = dup || id[k] == id[j];
    if (dup) id[j] = -1;
  }
  u64 run = 0;                                      // four 16-bit counts (a unit holds < 65536 pairs of a part)
#pragma unroll
  for (int j = 0; j < DMAX; j++)
    if (id[j] >= 0) run += 1ULL << (16 * part(id[j]));
  // inclusive scan over the unit's rows: within the wave, then the waves before
  u64 inc = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wtot[wave] = inc;
  __syncthreads();
  u64 tot = 0;
  for (int i = 0; i < PG_UNIT / 64; i++) {
    const u64 v = wtot[i];
    if (i < wave) inc += v;
    tot += v;
  }
  __syncthreads();                                  // wtot is free for the next unit
  u32* ul = list + (int64_t)u * PG_UNIT * sp;
  int pos[PG_MAXP], base = 0;
#pragma unroll
  for (int p = 0; p < PG_MAXP; p++) {
    pos[p] = base + (int)((inc >> (16 * p)) & 0xffff) - (int)((run >> (16 * p)) & 0xffff);
    if (t == 0 && p < P) poff[u * (PG_MAXP + 1) + p] = base;
    base += p < P ? (int)((tot >> (16 * p)) & 0xffff) : 0;
  }
  if (t == 0) poff[u * (PG_MAXP + 1) + P] = base;
#pragma unroll
  for (int j = 0; j < DMAX; j++)
    if (id[j] >= 0) {
      const int q = part(id[j]);
      int at = 0;
#pragma unroll
      for (int p = 0; p < PG_MAXP; p++)
        if (p == q) at = pos[p]++;
      ul[at] = ((u32)t << 22) | (u32)id[j];
    }
}

// every unit (workgroup u builds unit u)
template <int DMAX>
__global__ __launch_bounds__(PG_UNIT) void k_plist(const int* __restrict__ nbr, int n, int dpad, int sp, int P,
                                                   u32* __restrict__ list, int* __restrict__ poff)
{
  PlMove none;
  none.cnt = 0;
  plist_unit<DMAX>(nbr, n, dpad, sp, P, list, poff, blockIdx.x, none);
}

// Entries at positions the host found (a 2-opt move, checked with k_swap's
// rule on the host's copy of the adjacency, or the entries an upload changed):
// workgroup 0 stores the mv.cnt entries into the adjacency and the tables;
// workgroup i < units.cnt rebuilds the pair lists of unit units.u[i],
// applying the entries to its rows as it reads them.  With mv.cnt = 0 it only
// rebuilds the units.
struct PlUnits { int u[PL_MOVE_MAX]; int cnt; };
struct PlArgs { const int* nbr; int n, dpad, sp, P; u32* list; int* poff; };

template <int DMAX>
__global__ __launch_bounds__(PG_UNIT) void k_swap_at(int* adj, int degree, SwapTables tb, PlMove mv, PlUnits units,
                                                     PlArgs pl)
{
  if (blockIdx.x == 0 && (int)threadIdx.x < mv.cnt) {
    const int e = threadIdx.x;
    adj[(int64_t)mv.row[e] * degree + mv.pos[e]] = mv.to[e];
    const int64_t k = (int64_t)mv.row[e] * tb.tdp + mv.pos[e];
    if (tb.t16 && tb.t16[k] == (unsigned short)mv.from[e]) tb.t16[k] = (unsigned short)mv.to[e];
    if (tb.t32 && tb.t32[k] == mv.from[e]) tb.t32[k] = mv.to[e];
  }
  if ((int)blockIdx.x < units.cnt)
    plist_unit<DMAX>(pl.nbr, pl.n, pl.dpad, pl.sp, pl.P, pl.list, pl.poff, units.u[blockIdx.x], mv);
}

typedef u32 u32x4v __attribute__((ext_vector_type(4)));
typedef u32 u32x2v __attribute__((ext_vector_type(2)));

// Persistent: workgroup b runs on XCD b % 8 and takes every (gridDim.x / 8)-th
// unit of its XCD's contiguous range of units (a unit: PG_UNIT rows of one
// task), so the XCD's workgroups sweep that range together; any grid that is a
// multiple of 8 works.  The gathers are buffer loads with cache policy AUX
// (0: default; 1: sc0, 16: sc1, 2: nt), bounded by the slab (out-of-range
// offsets read zero).
template <bool LS, int AUX>
__global__ __launch_bounds__(PG_WG) void k_pgather(const LevelTask* __restrict__ tasks, int level, int ring,
                                                   int mode, LvList lst, PgArgs a)
{
  __shared__ u64x2 acc[PG_UNIT * 8];                  // row r of the unit at [8 r, 8 r + 8)
  __shared__ u64 s_pop[PG_WG / 64];
  const int x = blockIdx.x % ODP_NXCD, wi = blockIdx.x / ODP_NXCD, nwx = gridDim.x / ODP_NXCD;
  const long long u_lo = a.units * x / ODP_NXCD, u_hi = a.units * (x + 1) / ODP_NXCD;
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int nw = PG_WG / 64;
  const bool count = (mode & LV_COUNT) != 0, store = (mode & LV_STORE) != 0;
  const int w = 2 * (lane & 7);                       // the lane's two words of a row
  for (long long u = u_lo + wi; u < u_hi; u += nwx) { // uniform
    const int t = (int)(u / a.upt), ur = (int)(u - (long long)t * a.upt);
    const LevelTask& T = tasks[t];
    if (*T.done) continue;
    const u64 c1 = T.cnt[(int64_t)((level - 1) % ring) * T.cstride];
    if (c1 == (u64)T.n * (u64)T.l || (level >= 2 && c1 == T.cnt[(int64_t)((level - 2) % ring) * T.cstride])) {
      if (tid == 0) *T.done = 1;
      continue;
    }
    if (count && ur == 0 && tid == 0) T.cnt[(int64_t)((level + 1) % ring) * T.cstride] = 0;
    const u64* __restrict__ cur = (level & 1) ? T.buf1 : T.buf0;
    u64* __restrict__ nxt = (level & 1) ? T.buf0 : T.buf1;
    const __amdgpu_buffer_rsrc_t src = __builtin_amdgcn_make_buffer_rsrc((void*)cur, 0, a.n * 128, 0x00020000);
    const u64 f0 = full_word(T.l, w), f1 = full_word(T.l, w + 1);
#pragma unroll
    for (int i = tid; i < PG_UNIT * 8; i += PG_WG) { acc[i].x = 0; acc[i].y = 0; }
    __syncthreads();
    const u32* __restrict__ ul = a.list + (int64_t)ur * PG_UNIT * a.sp;
    const int* po = a.poff + ur * (PG_MAXP + 1);
    // the parts in order; within a part the waves take chunks of PG_CH pairs in
    // turn.  A chunk's PG_SLOTS gathers (eight rows each) are in flight while
    // the wave ORs in the previous chunk's rows: a slot's next gather is issued
    // as soon as its rows are ORed in, and a chunk's pairs are loaded a chunk
    // ahead.  (Waiting slot by slot instead of once per chunk was measured
    // slower.)  The in-flight rows live in registers: PG_SLOTS sets how many.
    const u32 SENT = 0xffffffffu;
    // the unit's part offsets, read once (no memory read in the chunk loop may
    // make the wave wait for the gathers in flight)
    const int b0 = __builtin_amdgcn_readfirstlane(po[0]), b1 = __builtin_amdgcn_readfirstlane(po[1]);
    const int b2 = a.P > 1 ? __builtin_amdgcn_readfirstlane(po[2]) : b1;
    const int b3 = a.P > 2 ? __builtin_amdgcn_readfirstlane(po[3]) : b2;
    const int b4 = a.P > 3 ? __builtin_amdgcn_readfirstlane(po[4]) : b3;
    // The XCD's workgroups take its units in rounds; every other round sweeps
    // the parts backwards, so it starts with the part the round before ended
    // with, whose rows are still in L2.
    const bool rev = (((u - u_lo - wi) / nwx) & 1) != 0;
    auto boff = [&](int p) { return p == 0 ? b0 : p == 1 ? b1 : p == 2 ? b2 : p == 3 ? b3 : b4; };
    auto pbeg = [&](int pp) { return boff(rev ? a.P - 1 - pp : pp); };
    auto pend = [&](int pp) { return boff(rev ? a.P - pp : pp + 1); };
    auto advance = [&](int& pp, int& cc) {                 // uniform
      while (pp < a.P && cc >= pend(pp)) { pp++; if (pp < a.P) cc = pbeg(pp) + wave * PG_CH; }
    };
    // a chunk's pairs, 64 per register (every lane loads: an idle one the unit's
    // first entry).  Lane group g (eight lanes, one row) takes the chunk's pairs
    // [g PG_SLOTS, (g + 1) PG_SLOTS), one per slot: a row's pairs of a part are
    // consecutive in the list, so the eight groups of a slot meet eight
    // different rows and their ORs no two lanes' same LDS word.  Register h,
    // lane l holds the pair of slot 8 h + l / 8, group l % 8.
    struct Pairs { u32 v[PG_NL]; };
    auto pairs = [&](int pp, int cc) -> Pairs {
      Pairs r;
#pragma unroll
      for (int h = 0; h < PG_NL; h++) {
        const int q = 8 * h + (lane >> 3);
        const int i = cc + (lane & 7) * PG_SLOTS + q;
        const bool ok = pp < a.P && q < PG_SLOTS && i < pend(pp);
        const u32 v = ul[ok ? i : 0];
        r.v[h] = ok ? v : SENT;
      }
      return r;
    };
    u64x2 gv[PG_SLOTS];
    int lr[PG_SLOTS];
    // Every lane loads (a pair-less lane past the slab's end, which reads zero):
    // loads under a lane mask would make the compiler wait for all of them.
    auto issue = [&](const Pairs& chunk, int q) {
      const u32 e = (u32)__shfl((int)chunk.v[q >> 3], (q & 7) * 8 + (lane >> 3), 64);
      lr[q] = e == SENT ? -1 : (int)(e >> 22);
      const u32 off = e == SENT ? 0xfffffff0u : (e & 0x3fffff) * 128u + w * 8u;
      const u32x4v v = __builtin_amdgcn_raw_buffer_load_b128(src, (int)off, 0, AUX);
      gv[q].x = ((u64)v.y << 32) | v.x; gv[q].y = ((u64)v.w << 32) | v.z;
    };
    int p0 = 0, k0 = pbeg(0) + wave * PG_CH;
    advance(p0, k0);
    int p1 = p0, k1 = k0 + nw * PG_CH;
    advance(p1, k1);
    const Pairs first = pairs(p0, k0);
    Pairs next = pairs(p1, k1);
#pragma unroll
    for (int q = 0; q < PG_SLOTS; q++) issue(first, q);
    while (p0 < a.P) {                                    // wave-uniform: this chunk's gathers are in flight
      int p2 = p1, k2 = k1 + nw * PG_CH;
      advance(p2, k2);
      const Pairs after = pairs(p2, k2);
      // Rows are 128 bytes apart: a row's eight lanes meet 16 of the 64 banks
      // with the first word of their pieces, the other 16 of the row's half of
      // the banks with the second.  A row with bit 1 set ORs its second word
      // first, so the rows of an instruction spread over four classes of 16
      // banks (row mod 4) instead of two.
#pragma unroll
      for (int q = 0; q < PG_SLOTS; q++) {
        if (lr[q] >= 0) {
          u64* d = (u64*)&acc[lr[q] * 8 + (w >> 1)];
          const int sw = (lr[q] >> 1) & 1;
          const u64 f = sw ? gv[q].y : gv[q].x, g = sw ? gv[q].x : gv[q].y;
          if (f) atomicOr(d + sw, f);
          if (g) atomicOr(d + (sw ^ 1), g);
        }
        issue(next, q);                                     // all SENT past the last chunk
      }
      p0 = p1; k0 = k1; p1 = p2; k1 = k2; next = after;
    }
    __syncthreads();
    // the unit's rows: written, counted, listed; wave w takes rows [R w, R w + R), R = PG_UNIT / nw
    u64 pop = 0;
    for (int g = 0; g < PG_UNIT / nw / 8; g++) {
      const int lr = wave * (PG_UNIT / nw) + g * 8 + (lane >> 3), r = ur * PG_UNIT + lr;
      const bool valid = r < a.n;
      const u64x2 v = acc[lr * 8 + (w >> 1)];
      if (valid) {
        if (store) {
          u64v2 o; o.x = v.x; o.y = v.y;
          __builtin_nontemporal_store(o, (u64v2*)(nxt + (int64_t)r * 16 + w));   // streamed past L2
        }
        pop += (u64)(__popcll(v.x) + __popcll(v.y));
      }
      if (LS) list_append(lst, valid && !(v.x == f0 && v.y == f1), ((u64)t << 48) | ((u64)r << 24) | (u64)w);
    }
    if (count) {
      pop = wave_sum(pop);
      if (lane == 0) s_pop[wave] = pop;
    }
    __syncthreads();                                  // acc and s_pop are free for the next unit
    if (count && tid == 0) {
      u64 sum = 0;
      for (int i = 0; i < nw; i++) sum += s_pop[i];
      if (sum) atomicAdd(&T.cnt[(int64_t)(level % ring) * T.cstride], sum);
    }
  }
}

// ---------------------------------------------------------------------------
// The BFS kind (the reference's CPU-only aspl_bfs, aspl.c:166-221, with the
// top-down step of utils.c:1999-2100): one breadth-first search per source
// over the row-neighbour table.  Workgroups are persistent and take sources
// from a counter; each keeps its visited bitmap in LDS (VIS_LDS) or in its own
// slice of device memory, and its two frontier queues (n ids each) in device
// memory.  A level's frontier entries (vertex, table position) are dealt over
// the threads; a test-and-set on the bitmap admits each new vertex once.  Per
// source the distances' sum, the eccentricity and the number reached are
// folded into per-workgroup totals, added into the result once at the end.
struct BfsOut {
  u64 sum;          // sum over the sources of sum_v dist(source, v)
  int ecc;          // largest eccentricity
  int disc;         // a source did not reach every vertex
};

__global__ void k_bfs_clear(BfsOut* out, int* next_src)
{
  if (threadIdx.x == 0) { out->sum = 0; out->ecc = 0; out->disc = 0; *next_src = 0; }
}

template <bool VIS_LDS>
__global__ __launch_bounds__(256) void k_bfs(const int* __restrict__ nbr, int n, int dpad, int nsrc,
                                             int* __restrict__ queues, u32* __restrict__ gvis,
                                             int* __restrict__ next_src, BfsOut* __restrict__ out)
{
  extern __shared__ u32 dyn_vis[];
  __shared__ int s_src, s_cnt;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int words = (n + 31) >> 5;
  u32* vis = VIS_LDS ? dyn_vis : gvis + (size_t)blockIdx.x * words;
  int* fq = queues + (size_t)blockIdx.x * 2 * n;
  int* nq = fq + n;
  u64 tot = 0;
  int ecc_max = 0, disc = 0;
  for (;;) {                                            // uniform: s_src is read after a barrier
    if (tid == 0) s_src = atomicAdd(next_src, 1);
    __syncthreads();
    const int src = __builtin_amdgcn_readfirstlane(s_src);
    if (src >= nsrc) break;
    for (int i = tid; i < words; i += nth) vis[i] = 0;
    __syncthreads();
    if (tid == 0) { vis[src >> 5] = 1u << (src & 31); fq[0] = src; }
    int fcount = 1, level = 0, reached = 1;
    u64 dsum = 0;
    while (fcount > 0) {                                // uniform
      level++;
      if (tid == 0) s_cnt = 0;
      __syncthreads();                                  // the frontier, the bitmap and s_cnt are set
      const int entries = fcount * dpad;
      for (int e = tid; e < entries; e += nth) {
        const int f = e / dpad, j = e - f * dpad;
        const int u = nbr[(size_t)fq[f] * dpad + j];
        const u32 b = 1u << (u & 31);
        if (!(atomicOr(&vis[u >> 5], b) & b)) nq[atomicAdd(&s_cnt, 1)] = u;
      }
      __syncthreads();                                  // every new vertex is queued and counted
      fcount = __builtin_amdgcn_readfirstlane(s_cnt);
      dsum += (u64)level * (u64)fcount;
      reached += fcount;
      int* t = fq; fq = nq; nq = t;
      __syncthreads();                                  // s_cnt is read before it is reset
    }
    tot += dsum;
    ecc_max = max(ecc_max, level - 1);
    if (reached < n) disc = 1;
  }
  if (tid == 0) {
    if (tot) atomicAdd(&out->sum, tot);
    atomicMax(&out->ecc, ecc_max);
    if (disc) atomicOr(&out->disc, 1);
  }
}

// ---------------------------------------------------------------------------
// k_sparse: the count of a level expected to be the last one (LV_SPARSE), from
// the list of chunks the level before it left unsaturated.  A chunk saturated
// at level L-1 keeps its bits at level L, so the count of level L is that of
// L-1 plus, over the listed chunks, the bits their neighbours' words add.  As
// for a count-only k_level launch nothing is written but the counts and the
// done flags.  When a sub-list overflowed its capacity every chunk of every
// running task is visited instead.  Any grid size works.

// the bits row r's 16-byte chunk at word w gains at this level (0 if saturated)
__device__ __forceinline__ int chunk_gain(const LevelTask& T, int level, int r, int w)
{
  const u64* __restrict__ cur = (level & 1) ? T.buf1 : T.buf0;
  const int64_t kwp = T.kwp;
  const u64 f0 = full_word(T.l, w), f1 = full_word(T.l, w + 1);
  const u64x2 x0 = *(const u64x2*)(cur + r * kwp + w);
  if (x0.x == f0 && x0.y == f1) return 0;
  u64x2 x = x0;
  const int* nb = T.nbr + (int64_t)r * T.dpad;
  for (int j = 0; j < T.dpad; j += 4) {            // dpad is a multiple of 4
    const int4 v = *(const int4*)(nb + j);
    const u64x2 a = *(const u64x2*)(cur + (int64_t)v.x * kwp + w);
    const u64x2 b = *(const u64x2*)(cur + (int64_t)v.y * kwp + w);
    const u64x2 c = *(const u64x2*)(cur + (int64_t)v.z * kwp + w);
    const u64x2 d = *(const u64x2*)(cur + (int64_t)v.w * kwp + w);
    x.x |= a.x | b.x | c.x | d.x;
    x.y |= a.y | b.y | c.y | d.y;
    if (x.x == f0 && x.y == f1) break;
  }
  return __popcll(x.x) + __popcll(x.y) - __popcll(x0.x) - __popcll(x0.y);
}

// adds each lane's gain to its task's count of the level: one wave sum and one
// atomic per task among the lanes with a gain (listed chunks next to each
// other in the list can come from several tasks; an atomic per lane on the
// few task counters had serialised, 15 ms for 3 M chunks at n = 20000).
// Called by whole waves.
__device__ __forceinline__ void add_gains(const LevelTask* __restrict__ tasks, int level, int ring,
                                          int t, int gain)
{
  bool left = gain != 0;
  for (u64 m = __ballot(left); m; m = __ballot(left)) {   // wave-uniform
    const int t0 = __shfl(t, __ffsll((long long)m) - 1, 64);
    const bool mine = left && t == t0;
    const u64 s = wave_sum(mine ? (u64)gain : 0ULL);
    if ((threadIdx.x & 63) == 0) atomicAdd(&tasks[t0].cnt[(int64_t)(level % ring) * tasks[t0].cstride], s);
    left = left && !mine;
  }
}

// the stop test of k_level at this level
__device__ __forceinline__ bool task_stops(const LevelTask& T, int level, int ring)
{
  const u64 c1 = T.cnt[(int64_t)((level - 1) % ring) * T.cstride];
  return c1 == (u64)T.n * (u64)T.l || (level >= 2 && c1 == T.cnt[(int64_t)((level - 2) % ring) * T.cstride]);
}

__global__ __launch_bounds__(LEVEL_WG) void k_sparse(const LevelTask* __restrict__ tasks, int ntasks,
                                                     int level, int ring, LvList lst)
{
  const u32 tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  // per running task: the count of level L starts from L-1's (slot L was
  // cleared by level L-1), and slot L+1 is cleared for level L+1
  for (u32 t = tid; t < (u32)ntasks; t += nth) {
    const LevelTask& T = tasks[t];
    if (*T.done) continue;
    if (task_stops(T, level, ring)) { *T.done = 1; continue; }
    T.cnt[(int64_t)((level + 1) % ring) * T.cstride] = 0;
    atomicAdd(&T.cnt[(int64_t)(level % ring) * T.cstride], T.cnt[(int64_t)((level - 1) % ring) * T.cstride]);
  }
  const u32 wbase = tid & ~63u;                       // the wave's first thread; loops are wave-uniform
  // lane s holds sub-list s's count (LIST_SUBS is the wave size) and the
  // count of the sub-lists before it: the sub-lists read as one list of
  // `total` entries, entry e in the last sub-list starting at or below e
  static_assert(LIST_SUBS == 64, "one sub-list per lane");
  const u32 cnt = lst.n[lane * LIST_PAD];
  if (!__any(cnt > lst.cap)) {
    u32 inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const u32 v = (u32)__shfl_up((int)inc, o, 64);
      if (lane >= o) inc += v;
    }
    const u32 excl = inc - cnt, total = (u32)__shfl((int)inc, 63, 64);
    // A task that stops at this level gains nothing: it stops because level
    // L-1 reached every pair (nothing listed) or added no bit (no chunk can
    // gain one now), so its listed chunks need no test.
    for (u32 b = wbase; b < total; b += nth) {
      const u32 e = b + lane;
      int sub = 0;                                  // every lane searches (the shuffles read all lanes)
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if ((u32)__shfl((int)excl, sub + step, 64) <= e) sub += step;
      const u32 at = e - (u32)__shfl((int)excl, sub, 64);
      int t = 0, gain = 0;
      if (e < total) {
        const u64 v = lst.e[(u64)sub * lst.cap + at];
        t = (int)(v >> 48);
        gain = chunk_gain(tasks[t], level, (int)((v >> 24) & 0xffffff), (int)(v & 0xffffff));
      }
      add_gains(tasks, level, ring, t, gain);
    }
  } else {
    for (int t = 0; t < ntasks; t++) {
      const LevelTask& T = tasks[t];
      if (*T.done || task_stops(T, level, ring)) continue;
      const int chunks = T.segs << T.lpr_log2;       // 16-byte chunks per row of the block
      const u64 total = (u64)T.n * (u64)chunks;
      for (u64 b = wbase; b < total; b += nth) {
        const u64 i = b + lane;
        int gain = 0;
        if (i < total) gain = chunk_gain(T, level, (int)(i / chunks), 2 * (int)(i % chunks));
        add_gains(tasks, level, ring, t, gain);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_tile: k levels of a grid graph per launch (temporal tiles)
//
// The table entries of a grid graph join cells at most lmax apart in x and in
// y, so a row's value k levels later depends only on the rows within k * lmax
// cells of it.  A workgroup loads one 16-byte chunk (128 sources, a 2-word
// column block) of a square tile of cells plus a ring of halo cells around it
// (tile + 2 * halo = 64 cells a side) into LDS, runs k <= halo / lmax levels
// there and writes the tile's rows back.  Level i of the launch updates the cells at least
// (i + 1) * lmax inside the ring: those are the cells later levels read, and
// every neighbour they read was updated at level i - 1.  Levels alternate
// between two LDS copies of the region (128 KiB, one barrier per level).
//
// A tile is skipped when it cannot change in this launch and its rows in the
// output buffer are already current: none of its rows changed in the previous
// launch, and none of its or its 8 neighbour tiles' rows changed at the
// previous launch's last level (a row changes at level L only if a row it
// gathers from changed at L-1, so changes spread at most lmax cells per
// level, and halo <= tile).  Counts are k_level's: level L+i's count is
// L-1's plus the gains of the tiles' rows up to level L+i.
#define TILE_RW 64                                // region side: tile + 2 * halo = one wave
#define TILE_SLOTS (TILE_RW * TILE_RW)
#define TILE_THREADS 1024
#define TILE_PER_THREAD (TILE_SLOTS / TILE_THREADS)
#define TILE_KMAX 8
#define TILE_DPMAX 16                             // widest table k_tile takes

struct TileArgs {
  const int* xy2row;     // cell x * height + y -> row
  const short* toff;     // [n][dpad]: entry j of row r as the cell offset dx * TILE_RW + dy (0: r itself)
  int width, height, tiles_x, tiles_y;
  int tile, halo;        // tile side and halo width in cells, tile + 2 * halo = TILE_RW, halo <= tile
  int lmax;              // largest |dx| or |dy| of an entry (>= 1)
  int k;                 // levels of this launch, k * lmax <= halo
  int dpar;              // the launch reads buf1 (1) or buf0 (0) and writes the other
  int fpar;              // parity of the tile flags it writes (it reads the other)
  int first;             // first launch of the evaluation: every tile runs
  int xcd;
};

// cell offsets of the row-neighbour table for k_tile; an entry longer than
// lmax (which the host measured on the adjacency) is an internal error
// One thread per cell.  A cell's offsets are stored in order of their y
// offset: k_tile's lanes are consecutive y of one region column and read
// neighbour j together, and a region slot's LDS banks follow its y (rows are
// TILE_RW slots apart), so lanes whose j-th neighbours share a y offset read
// distinct banks; random order put several lanes on one bank.
__global__ void k_tgeom(const int* __restrict__ nbr, int n, int dpad, const int* __restrict__ row2xy, int lmax,
                        short* __restrict__ toff, int* err)
{
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int a = row2xy[r];
    int o[TILE_DPMAX], key[TILE_DPMAX];
#pragma unroll
    for (int j = 0; j < TILE_DPMAX; j++) {
      o[j] = 0;
      key[j] = INT_MAX;                                      // past the width: sorts last
      if (j < dpad) {
        const int b = row2xy[nbr[(int64_t)r * dpad + j]];
        int dx = (b >> 16) - (a >> 16), dy = (b & 0xffff) - (a & 0xffff);
        if (dx > lmax || dx < -lmax || dy > lmax || dy < -lmax) { atomicOr(err, 8); dx = 0; dy = 0; }
        o[j] = dx * TILE_RW + dy;
        key[j] = dy;
      }
    }
    // odd-even transposition sort by y offset (every index fixed: registers only)
#pragma unroll
    for (int pass = 0; pass < TILE_DPMAX; pass++)
#pragma unroll
      for (int j = pass & 1; j + 1 < TILE_DPMAX; j += 2)
        if (key[j] > key[j + 1]) {
          const int tk = key[j]; key[j] = key[j + 1]; key[j + 1] = tk;
          const int to = o[j]; o[j] = o[j + 1]; o[j + 1] = to;
        }
#pragma unroll
    for (int j = 0; j < TILE_DPMAX; j++)
      if (j < dpad) toff[(int64_t)r * dpad + j] = (short)o[j];
  }
}

// k_tile_plan: per (task, tile) unit of the launch: k_level's stop test, the
// count slots of levels L..L+k-1 (tile 0 of a task), and the activity test;
// the active units go to a list that k_tile's workgroups drain.
// tq: [0..1] list lengths by flag parity, [2] queue head, [3..] the list.
__global__ void k_tile_plan(const LevelTask* __restrict__ tasks, int ntasks, int level, int ring, TileArgs g,
                            u32* __restrict__ tq)
{
  const int ntiles = g.tiles_x * g.tiles_y;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid == 0) { tq[g.fpar ^ 1] = 0; tq[2] = 0; }    // the next launch's list, this launch's queue
  for (int u = gid; u < ntasks * ntiles; u += gridDim.x * blockDim.x) {
    const int t = u / ntiles, tile = u - t * ntiles;
    const LevelTask& T = tasks[t];
    const int64_t cs = T.cstride;
    if (*T.done) continue;
    const u64 c1 = T.cnt[(int64_t)((level - 1) % ring) * cs];
    if (c1 == (u64)T.n * (u64)T.l || (level >= 2 && c1 == T.cnt[(int64_t)((level - 2) % ring) * cs])) {
      if (tile == 0) *T.done = 1;
      continue;
    }
    if (tile == 0) {
      // levels L..L+k-1 start from L-1's count; the next launch's slots are cleared
      for (int i = 0; i < g.k; i++) {
        atomicAdd(&T.cnt[(int64_t)((level + i) % ring) * cs], c1);
        T.cnt[(int64_t)((level + g.k + i) % ring) * cs] = 0;
        T.acnt[(int64_t)((level + g.k + i) % ring) * cs] = 0;
      }
    }
    // tile flags of the task: [parity][changed in the launch, changed at its last level][tile]
    const unsigned char* fr = T.chg0 + (g.fpar ^ 1) * 2 * ntiles;
    unsigned char* fw = T.chg0 + g.fpar * 2 * ntiles;
    const int tx = tile / g.tiles_y, ty = tile - tx * g.tiles_y;
    bool active = g.first != 0 || fr[tile] != 0;
    for (int ax = max(tx - 1, 0); ax <= min(tx + 1, g.tiles_x - 1) && !active; ax++)
      for (int ay = max(ty - 1, 0); ay <= min(ty + 1, g.tiles_y - 1) && !active; ay++)
        active = fr[ntiles + ax * g.tiles_y + ay] != 0;
    if (active) tq[3 + atomicAdd(&tq[g.fpar], 1u)] = (u32)u;
    else { fw[tile] = 0; fw[ntiles + tile] = 0; }
  }
}

// k_tile: persistent workgroups (one per CU: the two region copies take
// 128 KiB of LDS) pull units from the list until it is drained.  While a unit
// runs its levels, the workgroup already claims the next unit and looks up its
// cells' rows (after levels 0 and 1), so most of the chain of dependent global
// loads a unit starts with overlaps the previous unit's levels.

// cell (lx, ly) of the region of a tile: thread tid owns (wave + 16 m, lane),
// a wave whole region columns
struct TileUnit {
  int t, tile, tx, ty;
  __device__ TileUnit(u32 u, int ntiles, int tiles_y)
  {
    t = (int)(u / ntiles); tile = (int)(u - (u32)t * ntiles);
    tx = tile / tiles_y; ty = tile - tx * tiles_y;
  }
};

template <int DP>
__global__ __launch_bounds__(TILE_THREADS) void k_tile(const LevelTask* __restrict__ tasks, int level, int ring,
                                                       TileArgs g, u32* __restrict__ tq)
{
  __shared__ u64x2 rows[2][TILE_SLOTS];             // the region's chunks, cell (lx, ly) at lx * TILE_RW + ly
  __shared__ u64 red[TILE_THREADS / 64][TILE_KMAX]; // per wave and level: gains of the tile's rows
  __shared__ u32 claim_s;
  const int ntiles = g.tiles_x * g.tiles_y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const u32 nunits = tq[g.fpar];
  // rows of the unit's cells (-1: outside the grid) and their chunks
  auto cell_rows = [&](const TileUnit& U, int* rr) {
#pragma unroll
    for (int m = 0; m < TILE_PER_THREAD; m++) {
      const int x = U.tx * g.tile - g.halo + wave + m * (TILE_THREADS / 64), y = U.ty * g.tile - g.halo + lane;
      rr[m] = (x >= 0 && x < g.width && y >= 0 && y < g.height) ? g.xy2row[x * g.height + y] : -1;
    }
  };
  auto cell_chunks = [&](const TileUnit& U, const int* rr, u64x2* vv) {
    const LevelTask& T = tasks[U.t];
    const u64* __restrict__ in = g.dpar ? T.buf1 : T.buf0;
#pragma unroll
    for (int m = 0; m < TILE_PER_THREAD; m++) {
      vv[m].x = 0; vv[m].y = 0;
      if (rr[m] >= 0) vv[m] = *(const u64x2*)(in + (int64_t)rr[m] * T.kwp);
    }
  };
  if (tid == 0) claim_s = atomicAdd(&tq[2], 1u);
  __syncthreads();
  const u32 q0 = __builtin_amdgcn_readfirstlane(claim_s);
  if (q0 >= nunits) return;                          // uniform
  u32 u = tq[3 + q0];
  int rid[TILE_PER_THREAD];
  cell_rows(TileUnit(u, ntiles, g.tiles_y), rid);
  for (;;) {
    u32 claim = 0;
    if (tid == 0) claim = atomicAdd(&tq[2], 1u);    // the following unit, published after level 0
    const TileUnit U(u, ntiles, g.tiles_y);
    const LevelTask& T = tasks[U.t];
    const int64_t cs = T.cstride;
    unsigned char* fw = T.chg0 + g.fpar * 2 * ntiles;
    u64* __restrict__ out = g.dpar ? T.buf0 : T.buf1;
    const int64_t kwp = T.kwp;
    u32 off[TILE_PER_THREAD][DP / 2];
    const u64 f0 = full_word(T.l, 0), f1 = full_word(T.l, 1);   // a saturated chunk
    bool nzv = false, usv = false;                   // some cell of the thread's is nonzero / unsaturated
    {
      u64x2 val[TILE_PER_THREAD];
      cell_chunks(U, rid, val);
#pragma unroll
      for (int m = 0; m < TILE_PER_THREAD; m++) {
        rows[0][(wave + m * (TILE_THREADS / 64)) * TILE_RW + lane] = val[m];
        nzv = nzv || (val[m].x | val[m].y) != 0;
        usv = usv || (rid[m] >= 0 && !(val[m].x == f0 && val[m].y == f1));
      }
    }
#pragma unroll
    for (int m = 0; m < TILE_PER_THREAD; m++) {
#pragma unroll
      for (int j = 0; j < DP / 2; j++) off[m][j] = 0;
      if (rid[m] >= 0) {
        const uint2* o = (const uint2*)(g.toff + (int64_t)rid[m] * DP);
#pragma unroll
        for (int j = 0; j < DP / 4; j++) { const uint2 w = o[j]; off[m][2 * j] = w.x; off[m][2 * j + 1] = w.y; }
      }
    }
    // a region whose cells are all zero, or all saturated, stays so for the k
    // levels: its tile is written back as loaded without running them (most
    // units of a first launch, and tiles the BFS front has just left)
    const bool any_nz = __syncthreads_or(nzv) != 0;
    const bool live = any_nz && __syncthreads_or(usv) != 0;
    u32 qn = nunits, un = 0;
    int nrid[TILE_PER_THREAD];
#pragma unroll
    for (int m = 0; m < TILE_PER_THREAD; m++) nrid[m] = -1;
    if (!live) {
      if (tid == 0) claim_s = claim;
      __syncthreads();
      qn = __builtin_amdgcn_readfirstlane(claim_s);
      if (qn < nunits) un = tq[3 + qn];
    }
    for (int i = 0; i < (live ? g.k : 0); i++) {
      // level i reads copy i & 1 (level i - 1) and writes the other; cells it
      // leaves out are never read again
      const u64x2* __restrict__ src = rows[i & 1];
      u64x2* __restrict__ dst = rows[(i + 1) & 1];
      const int lo = (i + 1) * g.lmax, hi = TILE_RW - lo;
      u32 gain = 0;                                  // at most 4 * 128 bits a thread
#pragma unroll
      for (int m = 0; m < TILE_PER_THREAD; m++) {
        const int lx = wave + m * (TILE_THREADS / 64), sl = lx * TILE_RW + lane;
        if (rid[m] >= 0 && lx >= lo && lx < hi && lane >= lo && lane < hi) {
          const u64x2 o = src[sl];
          u64x2 x = o;
          // a saturated chunk gains nothing: its neighbours are not read
          if (!(o.x == f0 && o.y == f1))
#pragma unroll
          for (int j = 0; j < DP / 2; j++) {
            // the window keeps every neighbour inside the region: a cell at least
            // (i + 1) lmax from its edges, and no offset longer than lmax cells
            // (k_tgeom flags longer ones, which end the run)
            const int s0 = sl + (int)(short)(off[m][j] & 0xffff);
            const int s1 = sl + (int)(short)(off[m][j] >> 16);
            const u64x2 a = src[s0], c = src[s1];
            x.x |= a.x | c.x;
            x.y |= a.y | c.y;
          }
          dst[sl] = x;
          if (lx >= g.halo && lx < g.halo + g.tile && lane >= g.halo && lane < g.halo + g.tile)
            gain += (u32)(__popcll(x.x ^ o.x) + __popcll(x.y ^ o.y));   // x holds o: its new bits
        }
      }
      const u64 wgain = wave_sum((u64)gain);
      if (lane == 0) red[wave][i] = wgain;
      if (i == 0 && tid == 0) claim_s = claim;
      __syncthreads();                               // level i is in LDS
      // the next unit's loads, one dependent step per level
      if (i == 0) {
        qn = __builtin_amdgcn_readfirstlane(claim_s);
        if (qn < nunits) un = tq[3 + qn];
      } else if (i == 1) {
        if (qn < nunits) cell_rows(TileUnit(un, ntiles, g.tiles_y), nrid);
      }
    }
    if ((!live || g.k < 2) && qn < nunits) cell_rows(TileUnit(un, ntiles, g.tiles_y), nrid);   // not yet
    // the tile's rows, its counts and flags
#pragma unroll
    for (int m = 0; m < TILE_PER_THREAD; m++) {
      const int lx = wave + m * (TILE_THREADS / 64);
      if (rid[m] >= 0 && lx >= g.halo && lx < g.halo + g.tile && lane >= g.halo && lane < g.halo + g.tile)
        *(u64x2*)(out + (int64_t)rid[m] * kwp) = rows[live ? (g.k & 1) : 0][lx * TILE_RW + lane];
    }
    if (tid < g.k) {
      u64 cum = 0, last = 0;
      for (int i = 0; live && i <= tid; i++)
        for (int w = 0; w < TILE_THREADS / 64; w++) { cum += red[w][i]; if (i == tid) last += red[w][i]; }
      if (cum) atomicAdd(&T.cnt[(int64_t)((level + tid) % ring) * cs], cum);
      if (tid == g.k - 1) {
        fw[U.tile] = cum != 0;
        fw[ntiles + U.tile] = last != 0;
        const int cx = min(g.tile, g.width - U.tx * g.tile), cy = min(g.tile, g.height - U.ty * g.tile);
        atomicAdd(&T.acnt[(int64_t)(level % ring) * cs], (u64)(cx * cy));   // rows processed (timing)
      }
    }
    __syncthreads();                                 // red and the region are free for the next unit
    if (qn >= nunits) break;                         // uniform
    u = un;
#pragma unroll
    for (int m = 0; m < TILE_PER_THREAD; m++) rid[m] = nrid[m];
  }
}

// ---------------------------------------------------------------------------
// direct kernel: the first K levels of a group of column blocks in one launch
//
// Row v of level L holds the sources reachable from v by a walk of at most
// L+1 steps through the row-neighbour table (level 0: v and its neighbours;
// each level ORs in the neighbours' rows).  While those sets are small they
// are cheaper to enumerate than to gather.  One wave per row expands the walk
// tree of v breadth-first: the ids at depth d-1 sit in an LDS list, the wave
// loads their table rows (16 B per load, DIRECT_UNROLL independent loads per
// lane in flight), sets the bits in an LDS bitmap of the group's row and, for
// d < K, appends the ids to the depth-d list.  After each depth the bitmap is
// popcounted per block (the count of level d-1); after depth K the row is
// written to the buffer level K reads.  The sets are exactly those the level
// kernels build, so every count and every word equals k_level0 followed by
// levels 1..K-1.

struct DirectArgs {
  const int* nbr;        // [n][dpad] row-neighbour table
  const unsigned short* nbr16;   // the same with 16-bit ids (n <= 65536), read when ID16
  u64* out;              // block t, row v, word w: out[t * tstride + v * kw + w]
  int64_t tstride;       // words between the slabs of consecutive blocks
  u64* cnt;              // count of block t at level L: cnt[L * cstride + t], L < K
  int cstride;
  int n, dpad, kw, kwl;  // kw: words per block row, a power of two in [2, 128]; kwl = log2(kw)
  int src_lo, nsrc;      // sources of the group: [src_lo, src_lo + nsrc)
  int ntask, words;      // blocks in the group; words = ntask * kw
  int chunks;            // ceil(words / 128): 128-word chunks a wave sweeps per row
  int K;                 // levels computed (2..DIRECT_KMAX)
  int list;              // ints of walk-tree ids kept per wave: 4 + sum of dpad^i, 0 < i < K
  unsigned qmul;         // floor(2^32 / q) + 1 for q = 16-byte loads per table row > 1:
                         // idx / q == umulhi(idx, qmul) for idx < 2^32 / q (idx < 2^24, q <= 256 here)
  int pw[8];             // dpad^i
  int nt;                // the rows of level K are streamed past L2 (the next level gathers from other XCDs)
};

#define DIRECT_KMAX 6
#define DIRECT_UNROLL 8

// ALL: every id is a source of the group (src_lo == 0, nsrc == n), no range test
template <bool ALL>
__device__ __forceinline__ void direct_set(u32* bm, int id, int src_lo, int nsrc)
{
  const int c = id - src_lo;
  if (ALL || (unsigned)c < (unsigned)nsrc) atomicOr(&bm[c >> 5], 1u << (c & 31));
}

// ID16: the 16-bit table (eight ids per 16-byte load) instead of the 32-bit one
template <bool ALL, bool ID16>
__global__ __launch_bounds__(512) void k_direct(DirectArgs a)
{
  extern __shared__ __attribute__((aligned(16))) u64 dlds[];
  const int nw = blockDim.x >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  u64* bm = dlds + (int64_t)wave * a.words;
  u32* bm32 = (u32*)bm;
  // LDS: nw row bitmaps, nw id lists, then per-lane counts lc[K][chunks][64] of
  // the workgroup (lane l of chunk j holds words 128 j + 2 l, 128 j + 2 l + 1).
  // Depth 0 takes 4 ints of a list and depth d >= 1 dpad^d, so every depth
  // starts 16-byte aligned (dpad is a multiple of 4).
  int* list = (int*)(dlds + (int64_t)nw * a.words) + wave * a.list;
  u32* lc = (u32*)((int*)(dlds + (int64_t)nw * a.words) + nw * a.list);
  const int ncnt = a.K * a.chunks * 64;
  for (int i = threadIdx.x; i < ncnt; i += blockDim.x) lc[i] = 0;
  __syncthreads();
  const int IPL = ID16 ? 8 : 4;                    // ids per 16-byte load
  const int Q = a.dpad / IPL;                      // 16-byte loads per table row
  // Each wave works on its own rows, bitmap and lists: between the phases of a
  // row only the wave's own LDS accesses must be ordered, which LDS guarantees
  // for one wave's instructions, so a wavefront fence (compiler ordering) is
  // enough and the waves of a workgroup never wait for each other.
  for (int v = blockIdx.x * nw + wave; v < a.n; v += gridDim.x * nw) {   // wave-uniform
    for (int w = 2 * lane; w < a.words; w += 128) { u64x2 z; z.x = 0; z.y = 0; *(u64x2*)(bm + w) = z; }
    if (lane == 0) list[0] = v;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (lane == 0) direct_set<ALL>(bm32, v, a.src_lo, a.nsrc);
    int off = 0;                                   // start of the depth d-1 list
    for (int d = 1; d <= a.K; d++) {
      const int m = a.pw[d - 1];                   // ids at depth d-1
      const int jobs = m * Q;
      const int* from = list + off;
      const int step = d == 1 ? 4 : m;             // ints the depth d-1 list takes
      int* to = list + off + step;                 // the depth-d list follows
      for (int j0 = 0; j0 < jobs; j0 += 64 * DIRECT_UNROLL) {
        uint4 q[DIRECT_UNROLL];
#pragma unroll
        for (int u = 0; u < DIRECT_UNROLL; u++) {
          const unsigned idx = j0 + u * 64 + lane;
          if (idx < (unsigned)jobs) {
            const unsigned r = Q == 1 ? idx : __umulhi(idx, a.qmul);
            const int64_t e = (int64_t)from[r] * a.dpad + IPL * (idx - r * Q);
            q[u] = ID16 ? *(const uint4*)(a.nbr16 + e) : *(const uint4*)(a.nbr + e);
          }
        }
#pragma unroll
        for (int u = 0; u < DIRECT_UNROLL; u++) {
          const int idx = j0 + u * 64 + lane;
          if (idx < jobs) {
            if (ID16) {
              int4 lo, hi;
              lo.x = q[u].x & 0xffff; lo.y = q[u].x >> 16; lo.z = q[u].y & 0xffff; lo.w = q[u].y >> 16;
              hi.x = q[u].z & 0xffff; hi.y = q[u].z >> 16; hi.z = q[u].w & 0xffff; hi.w = q[u].w >> 16;
              direct_set<ALL>(bm32, lo.x, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, lo.y, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, lo.z, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, lo.w, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, hi.x, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, hi.y, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, hi.z, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, hi.w, a.src_lo, a.nsrc);
              if (d < a.K) { *(int4*)(to + 8 * idx) = lo; *(int4*)(to + 8 * idx + 4) = hi; }
            } else {
              int4 v;
              v.x = (int)q[u].x; v.y = (int)q[u].y; v.z = (int)q[u].z; v.w = (int)q[u].w;
              direct_set<ALL>(bm32, v.x, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, v.y, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, v.z, a.src_lo, a.nsrc);
              direct_set<ALL>(bm32, v.w, a.src_lo, a.nsrc);
              if (d < a.K) *(int4*)(to + 4 * idx) = v;
            }
          }
        }
      }
      off += step;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      // level d-1: popcounts into the lane's counters; after the last depth, the row itself
      u32* lcd = lc + (d - 1) * a.chunks * 64 + lane;
      for (int j = 0; j < a.chunks; j++) {
        const int w = 128 * j + 2 * lane;
        if (w < a.words) {
          const u64x2 x = *(const u64x2*)(bm + w);
          const u32 c = (u32)(__popcll(x.x) + __popcll(x.y));
          if (c) atomicAdd(&lcd[64 * j], c);
          if (d == a.K) {
            u64* dst = a.out + (int64_t)(w >> a.kwl) * a.tstride + (int64_t)v * a.kw + (w & (a.kw - 1));
            if (a.nt) {
              u64v2 o; o.x = x.x; o.y = x.y;
              __builtin_nontemporal_store(o, (u64v2*)dst);
            } else {
              *(u64x2*)dst = x;
            }
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
  }
  __syncthreads();
  // per block: the kw/2 lanes of a chunk that share a block row add up
  const int half = a.kw >> 1;
  for (int e = wave; e < a.K * a.chunks; e += nw) {            // wave-uniform
    const int L = e / a.chunks, j = e - L * a.chunks;
    u64 c = lc[e * 64 + lane];
    for (int o = 1; o < half; o <<= 1) c += __shfl_xor(c, o, 64);
    const int w = 128 * j + 2 * lane;
    if (w < a.words && (lane & (half - 1)) == 0 && c)
      atomicAdd(&a.cnt[(int64_t)L * a.cstride + (w >> a.kwl)], c);
  }
}

// ---------------------------------------------------------------------------
// resident kernel
//
// One workgroup runs every level of one block of sources with the block's
// bit matrix in LDS: u32[n][WB], a single buffer.  Each thread owns rows
// tid, tid+nth, ... (at most RPT of them) and keeps their next values in
// registers until the whole workgroup has read the level, then writes them
// back, so one buffer of n*WB*4 bytes suffices (two barriers per level).
// Neighbour ids are read as 16-bit row indices (graphs here have n <= 32768),
// eight per 16-byte load.

struct ResTask {
  const unsigned short* nbr;   // [n][dp] 16-bit row indices, padded with the row itself
  int n, dp;                   // rows; ids per row (multiple of 8)
  int src0, l;                 // first source row of the block, sources in it (<= 32*wb)
  int s;                       // symmetry factor of the graph
  int graph;                   // index into the per-graph outputs
  int wb;                      // 32-bit words per row of the block: 1, 2, 4 or 8
  int dq;                      // groups of 4 ids a row reads: ceil(table width / 4) <= dp / 4
  int deg;                     // table width (entries a row holds, padding included)
  int walk;                    // levels 0 .. walk - 1 from walks out of the block's sources (0: none)
};

// Levels a block builds from walks out of its l sources instead of pulling
// every row's neighbours: level L (sources within L + 1 hops) adds the walks
// of exactly L + 1 steps, l * deg^(L+1) LDS bit-sets (each decoded prefix of L
// steps reads one table row for its deg last steps), against n * (deg + 1) row
// reads for a pulled level; a level is walked while cost * (L + 1) times its
// walks stay below a pulled level's row reads (cost 1: measured best for the
// 1024-graph batch, 20.2 ms against 20.3 at 2, 20.55 at 4, 20.5 at 0.5).
// Returns the first pulled level (1: level 0 alone, whose l * (deg + 1) walks
// of at most one step always beat reading the table).
__host__ __device__ inline int res_walk_levels(int n, int deg, int l, double cost = 1.0)
{
  int K = 1;
  double w = (double)l * deg;                        // walks of one step
  for (int L = 1; L < 8; L++) {
    w *= deg;                                        // walks of L + 1 steps
    if (w * cost * (L + 1) > (double)n * (deg + 1)) break;
    K = L + 1;
  }
  return K;
}

struct ResOut {     // per-graph accumulators
  u64 sum;          // sum over blocks of sum_L (n*l - count_L) * s   (the n(n-1) base excluded)
  int diam;         // max over blocks of (terminal level + 1)
  u32 flags;        // bit 0: a block found no new bits before completing (disconnected)
  u32 levels;       // levels run, summed over the graph's blocks (for byte accounting)
  u32 pad;
};

// rows per thread held in registers for each width (at most 32 u32 of next rows)
// rows per thread of a resident block: with 4-word rows, 10 (10240 rows, as
// many as LDS holds: graphs of 8192 < n <= 10208 take 4-word blocks, 23.5 ms
// against 24.6 for the 1024-graph batch)
__host__ __device__ constexpr int res_rpt(int wb) { return wb == 1 ? 16 : wb == 2 ? 16 : wb == 4 ? 10 : 2; }
// rows per thread of the single-graph instances past 16384 rows, for two-word
// blocks (their LDS buffer holds up to 20416 rows) and one-word blocks (up to
// 40832 rows: 28 rows a thread without spills, 32 and 40 with 52 and 196 bytes
// a lane spilled, still as fast as the level plan or faster at every degree):
// instantiations of their own, so that the other launches keep their registers
#define RES_RPT_WIDE 20
#define RES_RPT_WIDE1 28
#define RES_RPT_WIDE1B 32
#define RES_RPT_WIDE1C 40

template <int WB> struct RowT { u32 w[WB]; };

template <int WB>
__device__ __forceinline__ RowT<WB> lds_row(const u32* p)
{
  RowT<WB> r;
  if constexpr (WB == 1) {
    r.w[0] = p[0];
  } else if constexpr (WB == 2) {
    const uint2 v = *(const uint2*)p;
    r.w[0] = v.x; r.w[1] = v.y;
  } else {
#pragma unroll
    for (int k = 0; k < WB; k += 4) {
      const uint4 v = *(const uint4*)(p + k);
      r.w[k] = v.x; r.w[k + 1] = v.y; r.w[k + 2] = v.z; r.w[k + 3] = v.w;
    }
  }
  return r;
}

template <int WB>
__device__ __forceinline__ void lds_store(u32* p, const RowT<WB>& r)
{
  if constexpr (WB == 1) {
    p[0] = r.w[0];
  } else if constexpr (WB == 2) {
    *(uint2*)p = make_uint2(r.w[0], r.w[1]);
  } else {
#pragma unroll
    for (int k = 0; k < WB; k += 4) *(uint4*)(p + k) = make_uint4(r.w[k], r.w[k + 1], r.w[k + 2], r.w[k + 3]);
  }
}

template <int WB>
__device__ __forceinline__ void row_bit(RowT<WB>& r, int c, int l)
{
  if ((unsigned)c < (unsigned)l) {
    const int k = c >> 5;
    const u32 bit = 1u << (c & 31);
#pragma unroll
    for (int i = 0; i < WB; i++) r.w[i] |= (i == k) ? bit : 0u;
  }
}

template <int WB>
__device__ __forceinline__ void row_or(RowT<WB>& r, const u32* cur, int v)
{
  const RowT<WB> a = lds_row<WB>(cur + v * WB);
#pragma unroll
  for (int k = 0; k < WB; k++) r.w[k] |= a.w[k];
}

// The two rows a 32-bit word of 16-bit ids names, ORed into r: their LDS byte
// addresses id * 4 WB + base each take one v_mad_u32_u16 (op_sel picks the
// high half), where the compiler's own extract, shift and add took two or
// three instructions -- the pulled levels are bound by VALU and LDS issue.
typedef __attribute__((address_space(3))) const u32 lds_u32_t;
template <int WB>
__device__ __forceinline__ RowT<WB> lds_row_at(u32 addr)
{
  typedef __attribute__((address_space(3))) const u32x2v lds_u32x2_t;
  typedef __attribute__((address_space(3))) const u32x4v lds_u32x4_t;
  RowT<WB> r;
  if constexpr (WB == 1) {
    r.w[0] = *(lds_u32_t*)(uintptr_t)addr;
  } else if constexpr (WB == 2) {
    const u32x2v v = *(lds_u32x2_t*)(uintptr_t)addr;
    r.w[0] = v.x; r.w[1] = v.y;
  } else {
#pragma unroll
    for (int k = 0; k < WB; k += 4) {
      const u32x4v v = *(lds_u32x4_t*)(uintptr_t)(addr + 4u * k);
      r.w[k] = v.x; r.w[k + 1] = v.y; r.w[k + 2] = v.z; r.w[k + 3] = v.w;
    }
  }
  return r;
}

template <int WB>
__device__ __forceinline__ void row_or2(RowT<WB>& r, u32 base, u32 word)
{
  u32 lo, hi;
  asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(lo) : "v"(word), "v"((u32)(4 * WB)), "v"(base));
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(hi) : "v"(word), "v"((u32)(4 * WB)), "v"(base));
  const RowT<WB> a = lds_row_at<WB>(lo), b = lds_row_at<WB>(hi);
#pragma unroll
  for (int k = 0; k < WB; k++) r.w[k] |= a.w[k] | b.w[k];
}

// No divergent branch may sit in a loop that contains a barrier.  With an
// `if (lane == 0)` around the queue pop, the compiler structurized the block
// loop per lane: lane 0 left the loop to pop the next block while the other
// lanes re-entered the barrier and re-read the old block id -- a livelock.  So
// every loop-control value is made wave-uniform with readfirstlane, and
// single-lane side effects are done by whole waves with zero contributions
// from the other lanes.
__device__ __forceinline__ u64 uniform_u64(u64 v)
{
  const u32 lo = __builtin_amdgcn_readfirstlane((u32)v);
  const u32 hi = __builtin_amdgcn_readfirstlane((u32)(v >> 32));
  return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 block_total(u64 v, u64* part, int wave, int nw)
{
  v = wave_sum(v);      // every lane holds the wave total: all write the same value
  part[wave] = v;
  __syncthreads();
  u64 s = 0;
  for (int i = 0; i < nw; i++) s += part[i];
  return uniform_u64(s);
}

// All levels of one block of T.l <= 32*WB sources; returns the block's totals
// (the same value in every thread).
// MAD: pulled rows address their neighbours with row_or2 (k_anneal keeps
// row_or: its larger register set spills more with it)
template <int WB, bool STOP = false, bool WALK = true, bool MAD = true, int RPT = res_rpt(WB)>
__device__ __forceinline__ ResOut res_run(const ResTask& T, u32* cur, u64 (*s_part)[16],
                                          int tid, int nth, int lane, int wave, int nw)
{
  const u64 target = (u64)T.n * (u64)T.l;
  const u32 cbase = (u32)(uintptr_t)(__attribute__((address_space(3))) u32*)cur;   // cur's LDS byte address
  RowT<WB> full;
#pragma unroll
  for (int k = 0; k < WB; k++) {
    const int v = T.l - 32 * k;
    full.w[k] = v >= 32 ? ~0u : (v <= 0 ? 0u : ((1u << v) - 1u));
  }
  // level 0 straight from the identity: row r holds r and its neighbours
  u64 pop = 0;
  if (WALK && T.walk >= 1) {
    // from the sources' side: clear every row, then each source sets its bit
    // in its own row and in its neighbours' rows (the table is symmetric)
    for (int r = tid; r < T.n; r += nth) {
      RowT<WB> row;
#pragma unroll
      for (int k = 0; k < WB; k++) row.w[k] = 0;
      lds_store<WB>(cur + r * WB, row);
    }
    __syncthreads();
    const int per = T.deg + 1;
    for (int it = tid; it < T.l * per; it += nth) {
      const int c0 = it / per, j = it - c0 * per;
      const int src = T.src0 + c0;
      const int v = j == T.deg ? src : (int)T.nbr[src * T.dp + j];
      atomicOr(&cur[v * WB + (c0 >> 5)], 1u << (c0 & 31));
    }
    __syncthreads();
    for (int r = tid; r < T.n; r += nth) {
      const RowT<WB> row = lds_row<WB>(cur + r * WB);
#pragma unroll
      for (int k = 0; k < WB; k++) pop += (u64)__popc(row.w[k]);
    }
  } else
  for (int r = tid; r < T.n; r += nth) {
    RowT<WB> row;
#pragma unroll
    for (int k = 0; k < WB; k++) row.w[k] = 0;
    row_bit<WB>(row, r - T.src0, T.l);
    // 8 ids per 16-byte load, then 4 more when the table width leaves them
    // (the padding past the width is not read)
    const uint4* ids = (const uint4*)T.nbr;     // uniform base, 32-bit offsets
    const u32 off = (u32)r * (u32)(T.dp >> 3);
    for (int j = 0; j < (T.dq >> 1); j++) {
      const uint4 v = ids[off + j];
      row_bit<WB>(row, (int)(v.x & 0xffff) - T.src0, T.l); row_bit<WB>(row, (int)(v.x >> 16) - T.src0, T.l);
      row_bit<WB>(row, (int)(v.y & 0xffff) - T.src0, T.l); row_bit<WB>(row, (int)(v.y >> 16) - T.src0, T.l);
      row_bit<WB>(row, (int)(v.z & 0xffff) - T.src0, T.l); row_bit<WB>(row, (int)(v.z >> 16) - T.src0, T.l);
      row_bit<WB>(row, (int)(v.w & 0xffff) - T.src0, T.l); row_bit<WB>(row, (int)(v.w >> 16) - T.src0, T.l);
    }
    if (T.dq & 1) {
      const uint2 v = ((const uint2*)T.nbr)[2 * off + (T.dq - 1)];
      row_bit<WB>(row, (int)(v.x & 0xffff) - T.src0, T.l); row_bit<WB>(row, (int)(v.x >> 16) - T.src0, T.l);
      row_bit<WB>(row, (int)(v.y & 0xffff) - T.src0, T.l); row_bit<WB>(row, (int)(v.y >> 16) - T.src0, T.l);
    }
    lds_store<WB>(cur + r * WB, row);
#pragma unroll
    for (int k = 0; k < WB; k++) pop += (u64)__popc(row.w[k]);
  }
  int par = 0;
  u64 c = block_total(pop, s_part[par], wave, nw);
  u64 prev = 0, sum = 0;
  int L = 0, D = 0;
  bool disc = false;
  // levels 1 .. T.walk - 1: the walks of exactly L + 1 steps from each source
  // set its bit in the rows they end at (the rows already hold level L - 1,
  // every shorter walk); the exits are left to the pulled levels' loop below
  while (WALK && L + 1 < T.walk && c != target && !(L >= 1 && c == prev) && L + 1 < T.n) {
    sum += (target - c) * (u64)T.s;
    prev = c;
    L++;
    const u32 deg = (u32)T.deg, steps = (u32)L + 1;
    // a thread takes a walk's first steps - 1 steps (its prefix, decoded with
    // divisions) and then every last step at once: the whole table row of the
    // vertex reached, 8 ids a load -- deg walks per decoded prefix
    u32 per = 1;
    for (u32 i = 0; i + 1 < steps; i++) per *= deg;
    const u32 prefixes = (u32)T.l * per;
    for (u32 it = (u32)tid; it < prefixes; it += (u32)nth) {
      const u32 c0 = it / per;
      u32 rem = it - c0 * per;
      int v = T.src0 + (int)c0;
      for (u32 i = 0; i + 1 < steps; i++) {
        const u32 j = rem % deg;
        rem /= deg;
        v = T.nbr[v * T.dp + (int)j];
      }
      u32* col = cur + (int)(c0 >> 5);
      const u32 b = 1u << (c0 & 31);
      const uint4* ids = (const uint4*)(T.nbr + v * T.dp);
      for (int q = 0; q < (T.dq >> 1); q++) {
        const uint4 g = ids[q];
        atomicOr(col + (g.x & 0xffff) * WB, b); atomicOr(col + (g.x >> 16) * WB, b);
        atomicOr(col + (g.y & 0xffff) * WB, b); atomicOr(col + (g.y >> 16) * WB, b);
        atomicOr(col + (g.z & 0xffff) * WB, b); atomicOr(col + (g.z >> 16) * WB, b);
        atomicOr(col + (g.w & 0xffff) * WB, b); atomicOr(col + (g.w >> 16) * WB, b);
      }
      if (T.dq & 1) {
        const uint2 g = ((const uint2*)ids)[T.dq - 1];
        atomicOr(col + (g.x & 0xffff) * WB, b); atomicOr(col + (g.x >> 16) * WB, b);
        atomicOr(col + (g.y & 0xffff) * WB, b); atomicOr(col + (g.y >> 16) * WB, b);
      }
    }
    __syncthreads();
    pop = 0;
    for (int r = tid; r < T.n; r += nth) {
      const RowT<WB> row = lds_row<WB>(cur + r * WB);
#pragma unroll
      for (int k = 0; k < WB; k++) pop += (u64)__popc(row.w[k]);
    }
    par ^= 1;
    c = block_total(pop, s_part[par], wave, nw);
  }
  for (;;) {
    if (c == target) { D = L + 1; break; }
    if ((L >= 1 && c == prev) || L + 1 >= T.n) { disc = true; break; }
    sum += (target - c) * (u64)T.s;
    prev = c;
    L++;
    RowT<WB> nx[RPT];
    pop = 0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int r = tid + i * nth;
      if (r < T.n) {
        RowT<WB> row = lds_row<WB>(cur + r * WB);
        bool sat = true;
#pragma unroll
        for (int k = 0; k < WB; k++) sat = sat && (row.w[k] == full.w[k]);
        if (!sat) {
          const uint4* ids = (const uint4*)T.nbr;
          const u32 off = (u32)r * (u32)(T.dp >> 3);
          bool done = false;
          for (int j = 0; j < (T.dq >> 1); j++) {
            const uint4 v = ids[off + j];
            if (MAD) {
              row_or2<WB>(row, cbase, v.x); row_or2<WB>(row, cbase, v.y);
              row_or2<WB>(row, cbase, v.z); row_or2<WB>(row, cbase, v.w);
            } else {
              row_or<WB>(row, cur, v.x & 0xffff); row_or<WB>(row, cur, v.x >> 16);
              row_or<WB>(row, cur, v.y & 0xffff); row_or<WB>(row, cur, v.y >> 16);
              row_or<WB>(row, cur, v.z & 0xffff); row_or<WB>(row, cur, v.z >> 16);
              row_or<WB>(row, cur, v.w & 0xffff); row_or<WB>(row, cur, v.w >> 16);
            }
            if (STOP) {                            // a row filled by 8 neighbours reads no more
              done = true;
#pragma unroll
              for (int k = 0; k < WB; k++) done = done && (row.w[k] == full.w[k]);
              if (done) break;
            }
          }
          if ((T.dq & 1) && !done) {
            const uint2 v = ((const uint2*)T.nbr)[2 * off + (T.dq - 1)];
            if (MAD) {
              row_or2<WB>(row, cbase, v.x); row_or2<WB>(row, cbase, v.y);
            } else {
              row_or<WB>(row, cur, v.x & 0xffff); row_or<WB>(row, cur, v.x >> 16);
              row_or<WB>(row, cur, v.y & 0xffff); row_or<WB>(row, cur, v.y >> 16);
            }
          }
        }
        nx[i] = row;
#pragma unroll
        for (int k = 0; k < WB; k++) pop += (u64)__popc(row.w[k]);
      }
    }
    par ^= 1;
    c = block_total(pop, s_part[par], wave, nw);   // every read of this level is done
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int r = tid + i * nth;
      if (r < T.n) lds_store<WB>(cur + r * WB, nx[i]);
    }
    __syncthreads();                               // the new level is visible
  }
  ResOut r;
  r.sum = sum; r.diam = D; r.flags = disc ? 1u : 0u; r.levels = (u32)(L + 1); r.pad = 0;
  return r;
}

// adds a block's totals into its graph's accumulators
__device__ __forceinline__ void res_commit(ResOut* out, int graph, const ResOut& b, int lane, int wave)
{
  if (wave == 0) {       // wave-uniform branch; only lane 0 contributes
    const bool first = lane == 0;
    atomicAdd(&out[graph].sum, first ? b.sum : 0ULL);
    atomicMax(&out[graph].diam, b.diam);
    atomicOr(&out[graph].flags, b.flags);
    atomicAdd(&out[graph].levels, first ? b.levels : 0u);
  }
}

// End of a batch of levels as the host reads it (pinned, coherent host memory).
struct LvHost {
  int err;          // the device error flag as it stood at the end of the batch
  u32 seq;          // the batch's sequence number, written after err and the counters
};

// The counters of levels [from, to) (ring slots L % ring of every task) and
// the error flag, stored into host memory; then the next sequence number
// (kept on the device, so that a replayed graph advances it too), which the
// host watches for instead of synchronising the stream.
__global__ __launch_bounds__(256) void k_lvfin(const u64* __restrict__ cnt, int ntasks, int from, int to, int ring,
                                              u64* hcnt, const int* err, LvHost* host, u32* dseq)
{
  const int m = (to - from) * ntasks;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const int L = from + i / ntasks, t = i - (i / ntasks) * ntasks;
    const int64_t k = (int64_t)(L % ring) * ntasks + t;
    hcnt[k] = cnt[k];
  }
  __threadfence_system();                          // each thread's stores reach host memory first
  __syncthreads();
  if (threadIdx.x == 0) {
    host->err = *err;
    const u32 sq = *dseq + 1;
    *dseq = sq;
    __threadfence_system();
    __hip_atomic_store(&host->seq, sq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Result of a single-graph launch as the host reads it (pinned, coherent host memory).
struct ResHost {
  ResOut out;
  int err;          // the device error flag as it stood at the end of the launch
  u32 seq;          // the launch's sequence number, written after out and err
};

// Finishing step of a single-graph launch.  Each workgroup stores its totals
// (over the blocks it ran) in its own slot; the last one to arrive adds the
// slots up, hands the result and the error flag to host memory and leaves the
// queue head and the arrival counter at zero for the next launch, so a launch
// needs no clear before it and no copy after it.
// Entries to change (2-opt moves): row, position, new id, applied in order to
// the 16-bit table and the adjacency of one graph.
#define RES_PATCH_MAX 8
struct ResPatch {
  int n;
  int e[RES_PATCH_MAX][3];
  unsigned short* table; int tdp;
  int* adj; int degree;
};

// one thread: in order, since a later entry may overwrite an earlier one
__device__ __forceinline__ void res_patch(const ResPatch& p, bool adj_too)
{
  for (int i = 0; i < p.n; i++) {
    const int r = p.e[i][0], c = p.e[i][1], val = p.e[i][2];
    p.table[(int64_t)r * p.tdp + c] = (unsigned short)val;
    if (adj_too) p.adj[(int64_t)r * p.degree + c] = val;
  }
}

__global__ void k_patch(ResPatch p)
{
  if (threadIdx.x == 0) res_patch(p, true);
}

struct ResFin {
  u32* done;           // workgroups finished; null: no finishing step (batched launches)
  ResOut* part;        // [gridDim.x] per-workgroup totals
  ResHost* host;       // device address of the host result
  const int* err;      // device error flag
  // applied before the levels: every workgroup stores the entries into the
  // table it reads (the same values: no conflict), workgroup 0 also into the
  // adjacency -- a move folded into the evaluation's launch
  ResPatch patch;
  u32 seq;             // stored into host->seq once the result is in host memory
};

__device__ __forceinline__ u32 wave_or(u32 v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// run by wave 0 of the last workgroup
__device__ __forceinline__ void res_finish(const ResFin& fin, int* qhead, int lane)
{
  __threadfence();
  u64 sum = 0, levels = 0;
  u32 flags = 0;
  int diam = 0;
  for (int i = lane; i < (int)gridDim.x; i += 64) {
    const ResOut* p = fin.part + i;
    sum += __hip_atomic_load(&p->sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    diam = max(diam, __hip_atomic_load(&p->diam, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    flags |= __hip_atomic_load(&p->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    levels += __hip_atomic_load(&p->levels, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  sum = wave_sum(sum);
  levels = wave_sum(levels);
  flags = wave_or(flags);
  diam = wave_max(diam);
  if (lane == 0) {
    ResOut o;
    o.sum = sum; o.diam = diam; o.flags = flags; o.levels = (u32)levels; o.pad = 0;
    fin.host->out = o;
    fin.host->err = __hip_atomic_load(fin.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence_system();                           // the result before its sequence number
    __hip_atomic_store(&fin.host->seq, fin.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(qhead, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(fin.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence_system();
  }
}

// The same loop for blocks of one width only: its register allocation is that
// width's alone (k_resident holds all four, and the widest pressure makes every
// width spill), and a launch can size its LDS to its own blocks.  STOP: a pulled
// row stops reading neighbours once full (measured: the 1024-graph batch 25.05
// -> 24.6 ms; one n=1024 graph 39.6k -> 36.0k evaluations/s, so only batches use it).
template <int WB, bool STOP = false, int RPT = res_rpt(WB)>
__global__ __launch_bounds__(1024) void k_resident_w(const ResTask* __restrict__ tasks, int ntasks,
                                                     int* __restrict__ qhead, ResOut* __restrict__ out, ResFin fin)
{
  extern __shared__ __attribute__((aligned(16))) u32 lds[];
  __shared__ int s_task, s_last;
  __shared__ u64 s_part[2][16];
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nth >> 6;
  ResOut mine;           // this workgroup's totals (single-graph launches)
  mine.sum = 0; mine.diam = 0; mine.flags = 0; mine.levels = 0; mine.pad = 0;
  if (fin.patch.n) {     // uniform: a kernel argument
    if (tid == 0) {
      res_patch(fin.patch, blockIdx.x == 0);
      __threadfence();   // stored at L2 before any of this workgroup's table loads
    }
    __syncthreads();
  }
  // the first gridDim.x blocks are dealt by workgroup index, the rest popped
  // from *qhead (which then counts past gridDim.x)
  int t = blockIdx.x;
  while (t < ntasks) {
    const ResOut b = res_run<WB, STOP, true, true, RPT>(tasks[t], lds, s_part, tid, nth, lane, wave, nw);
    if (fin.done) {
      mine.sum += b.sum; mine.diam = max(mine.diam, b.diam); mine.flags |= b.flags; mine.levels += b.levels;
    } else {
      res_commit(out, tasks[t].graph, b, lane, wave);
    }
    if ((int)gridDim.x >= ntasks) break;          // nothing left to pop
    if (wave == 0) {       // wave-uniform: lane 0 adds 1, the others add 0
      const int got = atomicAdd(qhead, lane == 0 ? 1 : 0);
      s_task = __builtin_amdgcn_readfirstlane(got);
    }
    __syncthreads();
    t = (int)gridDim.x + __builtin_amdgcn_readfirstlane(s_task);
  }
  if (fin.done) {        // uniform: a kernel argument
    if (tid == 0) {
      fin.part[blockIdx.x] = mine;
      __threadfence();   // the slot precedes the arrival
      s_last = atomicAdd(fin.done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && wave == 0) res_finish(fin, qhead, lane);
  }
}

// Persistent: every workgroup pulls blocks from *qhead until none is left, so
// the grid drains once the queue is exhausted.  Blocks of any width run in the
// same launch (T.wb picks the instantiation).  The launch's dynamic LDS must
// hold n * wb * 4 bytes for its largest block, and blockDim.x * res_rpt(wb)
// must cover n for every block.
__global__ __launch_bounds__(1024) void k_resident(const ResTask* __restrict__ tasks, int ntasks,
                                                   int* __restrict__ qhead, ResOut* __restrict__ out)
{
  extern __shared__ __attribute__((aligned(16))) u32 lds[];
  __shared__ int s_task;
  __shared__ u64 s_part[2][16];
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nth >> 6;
  for (;;) {
    if (wave == 0) {       // wave-uniform: lane 0 adds 1, the others add 0
      const int got = atomicAdd(qhead, lane == 0 ? 1 : 0);
      s_task = __builtin_amdgcn_readfirstlane(got);
    }
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_task);
    if (t >= ntasks) break;
    const ResTask T = tasks[t];
    ResOut b;
    switch (T.wb) {
      case 1: b = res_run<1>(T, lds, s_part, tid, nth, lane, wave, nw); break;
      case 2: b = res_run<2>(T, lds, s_part, tid, nth, lane, wave, nw); break;
      case 4: b = res_run<4>(T, lds, s_part, tid, nth, lane, wave, nw); break;
      default: b = res_run<8>(T, lds, s_part, tid, nth, lane, wave, nw); break;
    }
    res_commit(out, T.graph, b, lane, wave);
  }
}

// ---------------------------------------------------------------------------
// annealing on the device (k_anneal): many steps of the annealing loop of
// sample/SA/general.c:149-181 in one cooperative launch, for a regular general
// graph (s = 1) whose blocks fit LDS together with its neighbour table.
//
// Every workgroup holds the whole 16-bit neighbour table in LDS (the first
// `degree` entries of a row are its adjacency row) and evaluates one block of
// 32 * WB sources.  Thread 0 of every workgroup keeps the same copy of the
// chain's state -- generator, current and best results, temperature -- and
// draws the same 2-opt move the host's ODPX_Two_opt would draw from it, so the
// tables stay identical without any exchange.  One grid barrier per step: each
// workgroup publishes its block's totals, waits for all of them, adds them up
// in the same order and makes the same acceptance decision (the test of
// sample/SA/common.c:15-43); a rejected move is undone in every table.

struct AnnealState {
  uint64_t rng[4];
  double cur_a, best_a, temp, cooling, low_a;
  long long steps, accepted, best_sum;
  int cur_d, best_d, low_d;
  int optimum;          // the lower bounds were reached
  int stop;             // no valid move was found
  int hill, prio;
  int err;              // a grid barrier timed out
};

struct AnnealArgs {
  int* adj;             // [n][degree]: the chain's current graph (read at start, written at the end)
  int* best;            // [n][degree]: the best graph, written on every improvement
  AnnealState* st;
  ResOut* slots;        // [2][ntasks] block totals of a step (double buffered by step parity)
  u32* bar;             // grid barrier counter (zero at launch)
  int n, degree, dp, ntasks, steps;
};

#define ANNEAL_SPIN_LIMIT (1ull << 24)

__device__ __forceinline__ u32 anneal_rnd(uint64_t s[4], u32 bound)
{
  // xoshiro256** step and Lemire's multiply-shift, as odp_gen.c's rnd()
  const uint64_t r = ((s[1] * 5) << 7 | (s[1] * 5) >> 57) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = (s[3] << 45) | (s[3] >> 19);
  return (u32)(((r >> 32) * (uint64_t)bound) >> 32);
}

__device__ __forceinline__ uint64_t anneal_next(uint64_t s[4])
{
  const uint64_t r = ((s[1] * 5) << 7 | (s[1] * 5) >> 57) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = (s[3] << 45) | (s[3] >> 19);
  return r;
}

__device__ __forceinline__ bool anneal_has(const unsigned short* tab, int dp, int degree, int a, int b)
{
  for (int j = 0; j < degree; j++) if (tab[a * dp + j] == b) return true;
  return false;
}

__device__ __forceinline__ void anneal_replace(unsigned short* tab, int dp, int degree, int row, int from, int to)
{
  for (int j = 0; j < degree; j++) if (tab[row * dp + j] == from) { tab[row * dp + j] = (unsigned short)to; return; }
}

// a-b, c-d -> a-c, b-d in the table (ODPX_Swap_edges' order of replacements)
__device__ __forceinline__ void anneal_swap(unsigned short* tab, int dp, int degree, int a, int b, int c, int d)
{
  anneal_replace(tab, dp, degree, a, b, c);
  anneal_replace(tab, dp, degree, b, a, d);
  anneal_replace(tab, dp, degree, c, d, a);
  anneal_replace(tab, dp, degree, d, c, b);
}

template <int WB>
__global__ __launch_bounds__(1024) void k_anneal(AnnealArgs a)
{
  extern __shared__ __attribute__((aligned(16))) u32 lds[];
  __shared__ u64 s_part[2][16];
  __shared__ int s_mv[5], s_dec[3];
  __shared__ u64 s_sum;
  __shared__ int s_diam, s_flags;
  __shared__ AnnealState S;          // the chain's state; only thread 0 reads or writes it
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nth >> 6;
  const int g = blockIdx.x, G = gridDim.x, n = a.n, deg = a.degree, dp = a.dp;
  u32* bits = lds;
  unsigned short* tab = (unsigned short*)(lds + ((n * WB + 3) & ~3));
  for (int i = tid; i < n * dp; i += nth) {
    const int r = i / dp, j = i - r * dp;
    tab[i] = (unsigned short)(j < deg ? a.adj[(int64_t)r * deg + j] : r);
  }
  if (tid == 0) S = *a.st;
  __syncthreads();
  ResTask T;
  T.nbr = tab; T.n = n; T.dp = dp; T.src0 = g * 32 * WB; T.l = min(32 * WB, n - T.src0);
  T.s = 1; T.graph = 0; T.wb = WB; T.dq = (max(deg, 1) + 3) / 4;
  T.deg = deg; T.walk = 0;                            // res_run below is built without walks
  for (int k = 0; k < a.steps; k++) {                 // uniform: every exit below is uniform
    if (tid == 0) {
      int found = 0, ma = 0, mb = 0, mc = 0, md = 0;
      for (int tries = 0; tries < 1000000 && !found; tries++) {
        const int x = (int)anneal_rnd(S.rng, (u32)n), y = (int)anneal_rnd(S.rng, (u32)n);
        const int bx = tab[x * dp + anneal_rnd(S.rng, (u32)deg)];
        const int dy = tab[y * dp + anneal_rnd(S.rng, (u32)deg)];
        if (x == y || x == dy || bx == y || bx == dy) continue;
        if (anneal_has(tab, dp, deg, x, y) || anneal_has(tab, dp, deg, bx, dy)) continue;
        anneal_swap(tab, dp, deg, x, bx, y, dy);
        found = 1; ma = x; mb = bx; mc = y; md = dy;
      }
      s_mv[0] = found; s_mv[1] = ma; s_mv[2] = mb; s_mv[3] = mc; s_mv[4] = md;
    }
    __syncthreads();
    if (!s_mv[0]) {
      if (tid == 0) S.stop = 1;
      break;
    }
    const ResOut b = res_run<WB, false, false, false>(T, bits, s_part, tid, nth, lane, wave, nw);   // no walks (measured faster)
    const int par = k & 1;                             // slots of this step (k is uniform; S.steps is thread 0's)
    if (tid == 0) {
      a.slots[par * G + g] = b;
      __threadfence();
      atomicAdd(a.bar, 1u);
      const u32 target = (u32)G * (u32)(k + 1);
      unsigned long long spins = 0;
      int abort = 0;
      while (__hip_atomic_load(a.bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > ANNEAL_SPIN_LIMIT) { atomicOr(&a.st->err, 1); abort = 1; break; }
      }
      s_dec[2] = abort;
    }
    __syncthreads();
    if (s_dec[2]) break;
    // every block's totals, added in the same order in every workgroup
    if (wave == 0) {
      u64 sum = 0;
      int diam = 0;
      u32 flags = 0;
      for (int i = lane; i < G; i += 64) {
        const ResOut* p = a.slots + par * G + i;
        sum += __hip_atomic_load(&p->sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        diam = max(diam, __hip_atomic_load(&p->diam, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        flags |= __hip_atomic_load(&p->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      sum = wave_sum(sum); diam = wave_max(diam); flags = wave_or(flags);
      if (lane == 0) { s_sum = sum; s_diam = diam; s_flags = (int)flags; }
    }
    __syncthreads();
    if (tid == 0) {
      // the result as ODP_Set_aspl reports it (aspl.c:73-74, 332-336)
      long lsum = (long)n * (n - 1) + (long)s_sum;
      int d = max(1, s_diam);
      if (s_flags & 1) d = n + 1;
      double aspl = lsum / (((double)n - 1) * n);
      lsum /= 2.0;
      if (d > n) { d = INT_MAX; lsum = LONG_MAX; aspl = DBL_MAX; }
      S.steps++;
      int improved = 0, accept;
      if (d < S.best_d || (d == S.best_d && aspl < S.best_a)) {
        improved = 1;
        S.best_d = d; S.best_sum = lsum; S.best_a = aspl;
        if (d == S.low_d && aspl == S.low_a) S.optimum = 1;
      }
      if (S.optimum) accept = 1;              // the loop ends here; the move stays
      else if (S.hill) accept = aspl <= S.cur_a;
      else if (d < S.cur_d && !S.prio) accept = 1;
      else if (d > S.cur_d && !S.prio) accept = 0;
      else if (aspl <= S.cur_a) accept = 1;
      else {
        const double diff = (S.cur_a - aspl) * n * (n - 1);
        const double u = ((double)(anneal_next(S.rng) >> 11) + 0.5) / 9007199254740992.0;
        accept = exp(diff / S.temp) > u;
      }
      if (!S.optimum) {
        if (accept) { S.cur_d = d; S.cur_a = aspl; S.accepted++; }
        S.temp *= S.cooling;
      }
      s_dec[0] = improved; s_dec[1] = accept;
    }
    __syncthreads();
    if (s_dec[0] && g == 0)                            // the best graph: workgroup 0's table
      for (int i = tid; i < n * deg; i += nth) a.best[i] = tab[(i / deg) * dp + (i % deg)];
    __syncthreads();
    if (!s_dec[1] && tid == 0) anneal_swap(tab, dp, deg, s_mv[1], s_mv[3], s_mv[2], s_mv[4]);   // undo
    if (tid == 0) s_dec[2] = S.optimum;               // thread 0's state decides, for every thread
    __syncthreads();
    if (s_dec[2]) break;
  }
  __syncthreads();
  if (g == 0) {
    for (int i = tid; i < n * deg; i += nth) a.adj[i] = tab[(i / deg) * dp + (i % deg)];
    if (tid == 0) {
      S.err = __hip_atomic_load(&a.st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *a.st = S;
    }
  }
}
