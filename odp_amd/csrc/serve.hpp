// serve.hpp -- serving engine: continuous batching of 2-opt evaluation
// requests for many annealing chains whose graphs stay resident in HBM.
// Included once, at the end of aspl_hip.hip (same translation unit: it uses
// the Evaluator and the kernels).
//
// A chain (one graph) is registered once; its adjacency stays on the device,
// and small chains also keep their 16-bit neighbour table there.  A request
// is a list of 2-opt swaps followed by an exact evaluation.  The requests of
// one chain run one at a time in submission order (each is applied to the
// graph the previous one left).
//
// The engine is a progress engine driven by its caller: every ODPX_Serve_*
// call (submit, poll, wait, drain) first advances it -- collects the launches
// whose completion event has passed, then admits waiting requests into free
// slots -- so all HIP calls come from the calling thread, one call at a time
// (a mutex serialises callers).  Slots own a stream each, so their work
// overlaps on the device:
//   * a small slot runs every admissible small-chain request (up to
//     SV_MAX_BATCH) in ONE persistent k_resident launch;
//   * a large slot owns one set of level-plan bit buffers (an Evaluator) and
//     steps one large-chain request through its level batches.
// Requests that arrive while others are on the GPU are admitted as soon as a
// slot -- and with it its memory -- frees up.  Results never depend on what
// else is in flight: every request's integers come from its own blocks and
// counters.
#include <mutex>
#include <deque>
#include <memory>
#include <unordered_map>
#include <chrono>

static const int SV_MAX_BATCH = 64;       // requests per small-slot launch
// Gather levels a large slot enqueues per progress step.  Slots share HIP's
// hardware queues (4 per process), so a small slot's batch can sit behind a
// large slot's queued work; one level at a time bounds that wait to about one
// level (env ODP_HIP_SERVE_LEVELS overrides)
static int sv_large_levels()
{
  const char* v = getenv("ODP_HIP_SERVE_LEVELS");
  const int k = v ? atoi(v) : 1;
  return k >= 1 ? k : 1;
}

static double mono_time()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SwapJob {
  int* adj;                  // [n][degree]
  unsigned short* nbr16;     // [n][dp16] or null: mirrors adj in its first `degree` entries
  int n, degree, dp16, first, nswaps;
};

// One thread per chain; its swaps in order.  A swap applies only if a, b, c, d
// are distinct vertices and a-b, c-d are edges; otherwise it changes nothing
// and status[j] = 1.
__global__ void k_apply_swaps(const SwapJob* __restrict__ jobs, int njobs, const int* __restrict__ swaps,
                              int* __restrict__ status)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const SwapJob J = jobs[j];
  int bad = 0;
  for (int k = 0; k < J.nswaps; k++) {
    const int* q = swaps + 4 * (J.first + k);
    const int a = q[0], b = q[1], c = q[2], d = q[3];
    const int rows[4] = {a, b, c, d}, from[4] = {b, a, d, c}, to[4] = {c, d, a, b};
    bool ok = a != b && a != c && a != d && b != c && b != d && c != d;
    for (int e = 0; e < 4; e++) ok = ok && rows[e] >= 0 && rows[e] < J.n;
    int pos[4] = {-1, -1, -1, -1};
    for (int e = 0; ok && e < 4; e++) {
      const int* r = J.adj + (int64_t)rows[e] * J.degree;
      for (int i = 0; i < J.degree; i++)
        if (r[i] == from[e]) { pos[e] = i; break; }
      ok = pos[e] >= 0;
    }
    if (!ok) { bad = 1; continue; }
    for (int e = 0; e < 4; e++) {
      J.adj[(int64_t)rows[e] * J.degree + pos[e]] = to[e];
      if (J.nbr16) J.nbr16[(int64_t)rows[e] * J.dp16 + pos[e]] = (unsigned short)to[e];
    }
  }
  status[j] = bad;
}

struct SvChain {
  int n = 0, degree = 0, dp16 = 0, wb = 0;
  bool small = false;
  int* d_adj = nullptr;
  unsigned short* d_nbr16 = nullptr;
  size_t bytes = 0;
  std::deque<long long> queue;      // pending requests, FIFO
  bool busy = false;                // its front request is in a slot
};

struct SvReq {
  int chain = 0;
  std::vector<int> swaps;
  double t_submit = 0, t_start = 0, t_done = 0;
  int diameter = 0, status = 0;
  long sum = 0;
  double aspl = 0;
};

struct SvSmallSlot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<long long> ids;
  std::vector<int> rn;
  std::vector<ResTask> tasks;
  std::vector<SwapJob> jobs;
  std::vector<int> swaps;
  size_t t_cap = 0, s_cap = 0;
  ResTask* d_t = nullptr;
  SwapJob* d_jobs = nullptr;
  int *d_swaps = nullptr, *d_status = nullptr, *h_status = nullptr, *d_qhead = nullptr;
  ResOut *d_out = nullptr, *h_out = nullptr;
};

struct SvLargeSlot {
  Evaluator ev;                     // its stream is ev.st
  hipEvent_t done = nullptr;
  bool busy = false;
  long long id = 0;
  int chain = -1;
  int shaped_n = -1, shaped_d = -1;
  SwapJob *d_job = nullptr, *h_job = nullptr;    // h_job: pinned, read by the async copy
  int *d_swaps = nullptr, *d_status = nullptr, *h_status = nullptr;
  size_t s_cap = 0;
};

static struct Server {
  bool running = false;
  std::mutex mu;                               // one caller at a time drives the engine
  std::vector<SvChain> chains;
  std::unordered_map<long long, SvReq> reqs;   // submitted, not yet handed back
  std::deque<long long> completed;             // done, not yet polled
  long long next_id = 1;
  long long outstanding = 0;                   // submitted and not completed
  size_t budget = 0, used = 0;
  int nsmall = 0, nlarge = 0;                  // slots asked for at start
  std::vector<std::unique_ptr<SvSmallSlot>> small;
  std::vector<std::unique_ptr<SvLargeSlot>> large;
} S;

static bool sv_chain_ready(const SvChain& c, bool small) { return c.small == small && !c.busy && !c.queue.empty(); }

static void sv_complete(long long id)
{
  SvChain& c = S.chains[S.reqs[id].chain];
  c.busy = false;
  c.queue.pop_front();
  S.completed.push_back(id);
  S.outstanding--;
}

// true once the work recorded on `e` has finished
static bool sv_event_done(hipEvent_t e)
{
  const hipError_t q = hipEventQuery(e);
  if (q == hipSuccess) return true;
  if (q != hipErrorNotReady) HIPC(q);
  return false;
}

// ---- small slots: one k_resident launch over a batch of small chains ----

static void sv_small_launch(SvSmallSlot& sl)
{
  sl.ids.clear(); sl.rn.clear(); sl.jobs.clear(); sl.swaps.clear(); sl.tasks.clear();
  const double now = mono_time();
  size_t lds = 0;
  for (size_t i = 0; i < S.chains.size() && (int)sl.ids.size() < SV_MAX_BATCH; i++) {
    SvChain& c = S.chains[i];
    if (!sv_chain_ready(c, true)) continue;
    c.busy = true;
    const long long id = c.queue.front();
    SvReq& r = S.reqs[id];
    r.t_start = now;
    const int k = (int)sl.ids.size();
    sl.ids.push_back(id);
    SwapJob J;
    J.adj = c.d_adj; J.nbr16 = c.d_nbr16; J.n = c.n; J.degree = c.degree; J.dp16 = c.dp16;
    J.first = (int)(sl.swaps.size() / 4); J.nswaps = (int)(r.swaps.size() / 4);
    sl.swaps.insert(sl.swaps.end(), r.swaps.begin(), r.swaps.end());
    sl.jobs.push_back(J);
    sl.rn.push_back(c.n);
    const int b = 32 * c.wb;
    lds = std::max(lds, (size_t)c.n * c.wb * 4);
    for (int t = 0; t * b < c.n; t++) {
      ResTask T;
      T.nbr = c.d_nbr16; T.n = c.n; T.dp = c.dp16; T.src0 = t * b; T.l = std::min(b, c.n - t * b);
      T.s = 1; T.graph = k; T.wb = c.wb; T.dq = (std::max(c.degree, 1) + 3) / 4;
      T.deg = std::max(c.degree, 1); T.walk = res_walk_env(c.n, T.deg, T.l);
      sl.tasks.push_back(T);
    }
  }
  if (sl.ids.empty()) return;
  if (sl.tasks.size() > sl.t_cap) { dfree(&sl.d_t); dalloc(&sl.d_t, sl.tasks.size()); sl.t_cap = sl.tasks.size(); }
  if (sl.swaps.size() > sl.s_cap) { dfree(&sl.d_swaps); dalloc(&sl.d_swaps, sl.swaps.size()); sl.s_cap = sl.swaps.size(); }
  const int nb = (int)sl.ids.size();
  if (!sl.swaps.empty())
    HIPC(hipMemcpyAsync(sl.d_swaps, sl.swaps.data(), sizeof(int) * sl.swaps.size(), hipMemcpyHostToDevice, sl.st));
  HIPC(hipMemcpyAsync(sl.d_jobs, sl.jobs.data(), sizeof(SwapJob) * nb, hipMemcpyHostToDevice, sl.st));
  HIPC(hipMemcpyAsync(sl.d_t, sl.tasks.data(), sizeof(ResTask) * sl.tasks.size(), hipMemcpyHostToDevice, sl.st));
  hipLaunchKernelGGL(k_apply_swaps, dim3(1), dim3(SV_MAX_BATCH), 0, sl.st, sl.d_jobs, nb, sl.d_swaps, sl.d_status);
  HIPC(hipGetLastError());
  HIPC(hipMemsetAsync(sl.d_out, 0, sizeof(ResOut) * nb, sl.st));
  HIPC(hipMemsetAsync(sl.d_qhead, 0, sizeof(int), sl.st));
  const int grid = std::max(1, std::min((int)sl.tasks.size(), C.ncu));
  hipLaunchKernelGGL(k_resident, dim3(grid), dim3(1024), lds, sl.st, sl.d_t, (int)sl.tasks.size(), sl.d_qhead, sl.d_out);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(sl.h_out, sl.d_out, sizeof(ResOut) * nb, hipMemcpyDeviceToHost, sl.st));
  HIPC(hipMemcpyAsync(sl.h_status, sl.d_status, sizeof(int) * nb, hipMemcpyDeviceToHost, sl.st));
  HIPC(hipEventRecord(sl.done, sl.st));
  sl.busy = true;
}

static void sv_small_finish(SvSmallSlot& sl)
{
  const double now = mono_time();
  for (size_t k = 0; k < sl.ids.size(); k++) {
    SvReq& r = S.reqs[sl.ids[k]];
    const ResOut& o = sl.h_out[k];
    finish_one(sl.rn[k], o.sum, std::max(1, o.diam), (o.flags & 1) != 0, &r.diameter, &r.sum, &r.aspl);
    r.status = sl.h_status[k] ? -1 : 0;
    r.t_done = now;
    sv_complete(sl.ids[k]);
  }
  sl.busy = false;
}

// ---- large slots: one level-plan evaluation stepped through its batches ----

static void sv_large_launch(SvLargeSlot& sl, int ci)
{
  SvChain& c = S.chains[ci];
  c.busy = true;
  const long long id = c.queue.front();
  SvReq& r = S.reqs[id];
  r.t_start = mono_time();
  sl.id = id; sl.chain = ci;
  Evaluator& ev = sl.ev;
  if (!r.swaps.empty()) {
    if (r.swaps.size() > sl.s_cap) { dfree(&sl.d_swaps); dalloc(&sl.d_swaps, r.swaps.size()); sl.s_cap = r.swaps.size(); }
    SwapJob& J = *sl.h_job;
    J.adj = c.d_adj; J.nbr16 = nullptr; J.n = c.n; J.degree = c.degree; J.dp16 = 0; J.first = 0;
    J.nswaps = (int)(r.swaps.size() / 4);
    HIPC(hipMemcpyAsync(sl.d_swaps, r.swaps.data(), sizeof(int) * r.swaps.size(), hipMemcpyHostToDevice, ev.st));
    HIPC(hipMemcpyAsync(sl.d_job, sl.h_job, sizeof(SwapJob), hipMemcpyHostToDevice, ev.st));
    hipLaunchKernelGGL(k_apply_swaps, dim3(1), dim3(64), 0, ev.st, sl.d_job, 1, sl.d_swaps, sl.d_status);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(sl.h_status, sl.d_status, sizeof(int), hipMemcpyDeviceToHost, ev.st));
  } else {
    *sl.h_status = 0;
  }
  if (c.n != sl.shaped_n || c.degree != sl.shaped_d) {
    ev.shape(0, c.n, c.degree, 1, 0, -1, nullptr, 0);
    ODPX_Plan p{};
    p.plan = ODPX_PLAN_LEVEL; p.kw = LEVEL_BLOCK_WORDS; p.upw = 8; p.seg_major = 0; p.xcd = 1; p.group = 0;
    p.direct = -1; p.batch = 0;
    ev.apply_plan(p);
    sl.shaped_n = c.n; sl.shaped_d = c.degree;
  }
  ev.ext_adj = c.d_adj;
  ev.batch_cap = sv_large_levels();
  ev.lv_begin();
  HIPC(hipEventRecord(sl.done, ev.st));
  sl.busy = true;
}

// the slot's recorded work has finished: the next batch, or the result
static void sv_large_advance(SvLargeSlot& sl)
{
  Evaluator& ev = sl.ev;
  if (*ev.h_err == 0 && !ev.lv_step()) {
    HIPC(hipEventRecord(sl.done, ev.st));
    return;
  }
  SvReq& r = S.reqs[sl.id];
  if (*ev.h_err) {                 // cannot happen for a chain validated at registration
    *ev.h_err = 0;
    HIPC(hipMemsetAsync(ev.d_err, 0, sizeof(int), ev.st));
    HIPC(hipStreamSynchronize(ev.st));
    r.status = -2;
  } else {
    u64 excess; int d; bool disc;
    ev.lv_result(&excess, &d, &disc);
    long sm; double as;
    ev.result_of(excess, d, disc, &d, &sm, &as);
    if (d > ev.n) { d = INT_MAX; sm = LONG_MAX; as = DBL_MAX; }
    r.diameter = d; r.sum = sm; r.aspl = as;
    r.status = *sl.h_status ? -1 : 0;
  }
  r.t_done = mono_time();
  sv_complete(sl.id);
  sl.busy = false;
}

// collect finished work, then admit waiting requests into free slots
static void sv_progress()
{
  for (auto& p : S.small)
    if (p->busy && sv_event_done(p->done)) sv_small_finish(*p);
  for (auto& p : S.large)
    if (p->busy && sv_event_done(p->done)) sv_large_advance(*p);
  for (auto& p : S.small)
    if (!p->busy) sv_small_launch(*p);
  for (auto& p : S.large) {
    if (p->busy) continue;
    for (size_t i = 0; i < S.chains.size(); i++)
      if (sv_chain_ready(S.chains[i], false)) { sv_large_launch(*p, (int)i); break; }
  }
}

// blocks until some slot's recorded work has finished; false if none is busy
static bool sv_block()
{
  for (auto& p : S.small)
    if (p->busy) { HIPC(hipEventSynchronize(p->done)); return true; }
  for (auto& p : S.large)
    if (p->busy) { HIPC(hipEventSynchronize(p->done)); return true; }
  return false;
}

static void sv_small_free(SvSmallSlot& sl)
{
  dfree(&sl.d_t); dfree(&sl.d_jobs); dfree(&sl.d_swaps); dfree(&sl.d_status); hfree(&sl.h_status);
  dfree(&sl.d_out); hfree(&sl.h_out); dfree(&sl.d_qhead);
  HIPC(hipEventDestroy(sl.done));
  HIPC(hipStreamDestroy(sl.st));
}

static void sv_large_free(SvLargeSlot& sl)
{
  dfree(&sl.d_job); hfree(&sl.h_job); dfree(&sl.d_swaps); dfree(&sl.d_status); hfree(&sl.h_status);
  HIPC(hipEventDestroy(sl.done));
  hipStream_t st = sl.ev.st;
  sl.ev.release();
  HIPC(hipStreamDestroy(st));
}

// bytes of level-plan bit buffers one large slot holds for a chain of n vertices
static size_t sv_level_bytes(int n)
{
  int W = (n + 63) / 64;
  int kw = std::min(LEVEL_BLOCK_WORDS, W);
  size_t blocks = (W + kw - 1) / kw;
  return blocks * 2 * (size_t)n * kw * 8 + (size_t)n * round_up(64, 4) * 4;
}

extern "C" int ODPX_Serve_start(int small_slots, int large_slots, double hbm_budget_gb)
{
  RandFence rand_fence;
  ctx_init();
  std::lock_guard<std::mutex> lk(S.mu);
  if (S.running) return -1;
  S.budget = (size_t)(hbm_budget_gb * 1e9);
  S.used = 0;
  S.next_id = 1;
  S.outstanding = 0;
  S.nsmall = std::max(1, small_slots);
  S.nlarge = std::max(0, large_slots);
  for (int i = 0; i < S.nsmall; i++) {
    std::unique_ptr<SvSmallSlot> p(new SvSmallSlot());
    HIPC(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
    dalloc(&p->d_jobs, SV_MAX_BATCH); dalloc(&p->d_status, SV_MAX_BATCH); halloc(&p->h_status, SV_MAX_BATCH);
    dalloc(&p->d_out, SV_MAX_BATCH); halloc(&p->h_out, SV_MAX_BATCH); dalloc(&p->d_qhead, 1);
    S.small.push_back(std::move(p));
  }
  S.running = true;
  return 0;
}

// the large slots, created with the first large chain
static void sv_make_large_slots()
{
  for (int i = 0; i < S.nlarge; i++) {
    std::unique_ptr<SvLargeSlot> p(new SvLargeSlot());
    HIPC(hipStreamCreateWithFlags(&p->ev.st, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
    dalloc(&p->d_job, 1); halloc(&p->h_job, 1); dalloc(&p->d_status, 1); halloc(&p->h_status, 1);
    S.large.push_back(std::move(p));
  }
}

extern "C" int ODPX_Serve_add_chain(int nodes, int degree, const int* adjacency)
{
  RandFence rand_fence;
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.running || nodes < 2 || degree < 1 || !adjacency) return -1;
  for (long i = 0; i < (long)nodes * degree; i++)
    if (adjacency[i] < 0 || adjacency[i] >= nodes) return -1;     // every entry a vertex id
  SvChain c;
  c.n = nodes; c.degree = degree; c.dp16 = round_up(degree, 8);
  c.wb = batch_wb(nodes);
  c.small = c.wb > 0;
  c.bytes = (size_t)nodes * degree * 4 + (c.small ? (size_t)nodes * c.dp16 * 2 : 0);
  size_t extra = 0;
  if (!c.small) {
    // the first large chain of a size brings the large slots' buffers
    bool have = false;
    for (const SvChain& x : S.chains) if (!x.small && x.n >= nodes) have = true;
    if (!have && S.nlarge > 0) extra = (size_t)S.nlarge * sv_level_bytes(nodes);
  }
  if (S.used + c.bytes + extra > S.budget) return -1;      // admission by HBM footprint
  dalloc(&c.d_adj, (size_t)nodes * degree);
  HIPC(hipMemcpy(c.d_adj, adjacency, sizeof(int) * nodes * degree, hipMemcpyHostToDevice));
  if (c.small) {
    dalloc(&c.d_nbr16, (size_t)nodes * c.dp16);
    BatchGraph G;
    G.adj_off = 0; G.nbr_off = 0; G.ent_start = 0; G.n = nodes; G.degree = degree; G.dp = c.dp16; G.pad = 0;
    BatchGraph* d_g;
    int *d_err, h_err = 0;
    dalloc(&d_g, 1); dalloc(&d_err, 1);
    HIPC(hipMemcpy(d_g, &G, sizeof G, hipMemcpyHostToDevice));
    HIPC(hipMemset(d_err, 0, sizeof(int)));
    long long total = (long long)nodes * c.dp16;
    hipLaunchKernelGGL(k_prepare_batch, dim3((int)std::min<long long>((total + 255) / 256, 4096)), dim3(256), 0, 0,
                       d_g, 1, total, c.d_adj, c.d_nbr16, d_err);
    HIPC(hipGetLastError());
    HIPC(hipMemcpy(&h_err, d_err, sizeof(int), hipMemcpyDeviceToHost));
    dfree(&d_g); dfree(&d_err);
    if (h_err) { dfree(&c.d_adj); dfree(&c.d_nbr16); return -1; }
  }
  S.used += c.bytes + extra;
  if (!c.small && S.large.empty()) sv_make_large_slots();
  S.chains.push_back(c);
  return (int)S.chains.size() - 1;
}

extern "C" long long ODPX_Serve_submit(int chain, int nswaps, const int* swaps)
{
  RandFence rand_fence;
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.running || chain < 0 || chain >= (int)S.chains.size() || nswaps < 0) return -1;
  if (!S.chains[chain].small && S.nlarge == 0) return -1;
  const long long id = S.next_id++;
  SvReq r;
  r.chain = chain;
  if (nswaps) r.swaps.assign(swaps, swaps + 4 * nswaps);
  r.t_submit = mono_time();
  S.reqs[id] = r;
  S.chains[chain].queue.push_back(id);
  S.outstanding++;
  sv_progress();
  return id;
}

static void sv_fill(long long id, const SvReq& r, ODPX_Result* o)
{
  o->id = id; o->chain = r.chain; o->status = r.status;
  o->diameter = r.diameter; o->sum = r.sum; o->ASPL = r.aspl;
  o->t_submit = r.t_submit; o->t_start = r.t_start; o->t_done = r.t_done;
}

extern "C" int ODPX_Serve_poll(ODPX_Result* out, int max)
{
  RandFence rand_fence;
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.running) return 0;
  sv_progress();
  int k = 0;
  while (k < max && !S.completed.empty()) {
    long long id = S.completed.front();
    S.completed.pop_front();
    auto it = S.reqs.find(id);
    sv_fill(id, it->second, &out[k++]);
    S.reqs.erase(it);
  }
  return k;
}

extern "C" int ODPX_Serve_wait(long long id, ODPX_Result* out)
{
  RandFence rand_fence;
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.running || S.reqs.find(id) == S.reqs.end()) return -1;
  for (;;) {
    sv_progress();
    for (auto it = S.completed.begin(); it != S.completed.end(); ++it)
      if (*it == id) {
        S.completed.erase(it);
        auto r = S.reqs.find(id);
        sv_fill(id, r->second, out);
        S.reqs.erase(r);
        return 0;
      }
    if (!sv_block()) return -1;      // nothing in flight: the request cannot advance
  }
}

static void sv_drain_locked()
{
  for (;;) {
    sv_progress();
    if (S.outstanding == 0 || !sv_block()) return;
  }
}

extern "C" void ODPX_Serve_drain(void)
{
  RandFence rand_fence;
  std::lock_guard<std::mutex> lk(S.mu);
  if (S.running) sv_drain_locked();
}

extern "C" void ODPX_Serve_stop(void)
{
  RandFence rand_fence;
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.running) return;
  sv_drain_locked();
  for (auto& p : S.small) sv_small_free(*p);
  for (auto& p : S.large) sv_large_free(*p);
  S.small.clear();
  S.large.clear();
  for (SvChain& c : S.chains) { dfree(&c.d_adj); dfree(&c.d_nbr16); }
  S.chains.clear();
  S.reqs.clear();
  S.completed.clear();
  S.running = false;
}
