// serve.hpp -- serving engine: continuous batching of 2-opt evaluation
// requests for many annealing chains whose graphs stay resident in HBM.
// Included once, at the end of aspl_hip.hip (same translation unit: it uses
// the Evaluator and the kernels).
//
// A chain (one graph) is registered once; its adjacency stays on the device,
// and small chains also keep their 16-bit neighbour table there.  A request
// is a list of 2-opt swaps followed by an exact evaluation.  The requests of
// one chain run one at a time in submission order (each is applied to the
// graph the previous one left).  Worker threads, each with its own HIP stream,
// take work as soon as they are free:
//   * small workers pull every admissible small-chain request (up to
//     SV_MAX_BATCH) into ONE persistent k_resident launch;
//   * large workers own one set of level-plan bit buffers each (the HBM slot)
//     and run one large-chain request at a time.
// Requests that arrive while others are on the GPU are admitted the moment a
// worker -- and with it its slot and memory -- frees up; launches on different
// streams overlap on the device.  Results never depend on what else is in
// flight: every request's integers come from its own blocks and counters.
#include <thread>
#include <mutex>
#include <condition_variable>
#include <deque>
#include <unordered_map>
#include <chrono>

static const int SV_MAX_BATCH = 64;       // requests per small-worker launch

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
  bool busy = false;
};

struct SvReq {
  int chain = 0;
  std::vector<int> swaps;
  double t_submit = 0, t_start = 0, t_done = 0;
  int diameter = 0, status = 0;
  long sum = 0;
  double aspl = 0;
};

static struct Server {
  bool running = false, stopping = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<SvChain> chains;
  std::unordered_map<long long, SvReq> reqs;   // submitted, not yet handed back
  std::deque<long long> completed;             // done, not yet polled
  long long next_id = 1;
  long long outstanding = 0;                   // submitted and not completed
  size_t budget = 0, used = 0;
  int large_slots = 0;
  std::vector<std::thread> workers;
} S;

static bool sv_chain_ready(const SvChain& c, bool small) { return c.small == small && !c.busy && !c.queue.empty(); }

static void sv_complete(long long id, const SvReq& r)
{
  // caller holds S.mu
  SvChain& c = S.chains[r.chain];
  c.busy = false;
  c.queue.pop_front();
  S.completed.push_back(id);
  S.outstanding--;
}

static void sv_small_worker()
{
  hipStream_t st;
  HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  std::vector<ResTask> tasks;
  std::vector<SwapJob> jobs;
  std::vector<int> swaps;
  size_t t_cap = 0, s_cap = 0;
  ResTask* d_t = nullptr;
  SwapJob* d_jobs = nullptr;
  int *d_swaps = nullptr, *d_status = nullptr, *h_status = nullptr, *d_qhead = nullptr;
  ResOut *d_out = nullptr, *h_out = nullptr;
  dalloc(&d_jobs, SV_MAX_BATCH); dalloc(&d_status, SV_MAX_BATCH); halloc(&h_status, SV_MAX_BATCH);
  dalloc(&d_out, SV_MAX_BATCH); halloc(&h_out, SV_MAX_BATCH); dalloc(&d_qhead, 1);
  for (;;) {
    std::vector<long long> ids;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      S.cv_work.wait(lk, [] {
        if (S.stopping) return true;
        for (const SvChain& c : S.chains) if (sv_chain_ready(c, true)) return true;
        return false;
      });
      if (S.stopping) break;
      double now = mono_time();
      for (size_t i = 0; i < S.chains.size() && (int)ids.size() < SV_MAX_BATCH; i++) {
        SvChain& c = S.chains[i];
        if (!sv_chain_ready(c, true)) continue;
        c.busy = true;
        long long id = c.queue.front();
        S.reqs[id].t_start = now;
        ids.push_back(id);
      }
    }
    // snapshot what the launch needs (the request map is only mutated under the lock)
    jobs.clear(); swaps.clear(); tasks.clear();
    std::vector<int> rn(ids.size());
    size_t lds = 0;
    {
      std::lock_guard<std::mutex> lk(S.mu);
      for (size_t k = 0; k < ids.size(); k++) {
        const SvReq& r = S.reqs[ids[k]];
        const SvChain& c = S.chains[r.chain];
        SwapJob J;
        J.adj = c.d_adj; J.nbr16 = c.d_nbr16; J.n = c.n; J.degree = c.degree; J.dp16 = c.dp16;
        J.first = (int)(swaps.size() / 4); J.nswaps = (int)(r.swaps.size() / 4);
        swaps.insert(swaps.end(), r.swaps.begin(), r.swaps.end());
        jobs.push_back(J);
        rn[k] = c.n;
        const int b = 32 * c.wb;
        lds = std::max(lds, (size_t)c.n * c.wb * 4);
        for (int t = 0; t * b < c.n; t++) {
          ResTask T;
          T.nbr = c.d_nbr16; T.n = c.n; T.dp = c.dp16; T.src0 = t * b; T.l = std::min(b, c.n - t * b);
          T.s = 1; T.graph = (int)k; T.wb = c.wb; T.pad = 0;
          tasks.push_back(T);
        }
      }
    }
    if (tasks.size() > t_cap) { dfree(&d_t); dalloc(&d_t, tasks.size()); t_cap = tasks.size(); }
    if (swaps.size() > s_cap) { dfree(&d_swaps); dalloc(&d_swaps, std::max<size_t>(swaps.size(), 4)); s_cap = swaps.size(); }
    if (!swaps.empty())
      HIPC(hipMemcpyAsync(d_swaps, swaps.data(), sizeof(int) * swaps.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(SwapJob) * jobs.size(), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(d_t, tasks.data(), sizeof(ResTask) * tasks.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_apply_swaps, dim3(1), dim3(SV_MAX_BATCH), 0, st, d_jobs, (int)jobs.size(), d_swaps, d_status);
    HIPC(hipGetLastError());
    HIPC(hipMemsetAsync(d_out, 0, sizeof(ResOut) * ids.size(), st));
    HIPC(hipMemsetAsync(d_qhead, 0, sizeof(int), st));
    int grid = std::max(1, std::min((int)tasks.size(), C.ncu));
    hipLaunchKernelGGL(k_resident, dim3(grid), dim3(1024), lds, st, d_t, (int)tasks.size(), d_qhead, d_out);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(h_out, d_out, sizeof(ResOut) * ids.size(), hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(h_status, d_status, sizeof(int) * ids.size(), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    double now = mono_time();
    {
      std::lock_guard<std::mutex> lk(S.mu);
      for (size_t k = 0; k < ids.size(); k++) {
        SvReq& r = S.reqs[ids[k]];
        const ResOut& o = h_out[k];
        finish_one(rn[k], o.sum, std::max(1, o.diam), (o.flags & 1) != 0, &r.diameter, &r.sum, &r.aspl);
        r.status = h_status[k] ? -1 : 0;
        r.t_done = now;
        sv_complete(ids[k], r);
      }
    }
    S.cv_done.notify_all();
    S.cv_work.notify_all();
  }
  dfree(&d_t); dfree(&d_jobs); dfree(&d_swaps); dfree(&d_status); hfree(&h_status);
  dfree(&d_out); hfree(&h_out); dfree(&d_qhead);
  HIPC(hipStreamDestroy(st));
}

static void sv_large_worker()
{
  Evaluator ev;
  HIPC(hipStreamCreateWithFlags(&ev.st, hipStreamNonBlocking));
  SwapJob* d_job = nullptr;
  int *d_swaps = nullptr, *d_status = nullptr, *h_status = nullptr;
  size_t s_cap = 0;
  dalloc(&d_job, 1); dalloc(&d_status, 1); halloc(&h_status, 1);
  int shaped_n = -1, shaped_d = -1;
  for (;;) {
    long long id = 0;
    SvReq r;
    SvChain c;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      S.cv_work.wait(lk, [] {
        if (S.stopping) return true;
        for (const SvChain& x : S.chains) if (sv_chain_ready(x, false)) return true;
        return false;
      });
      if (S.stopping) break;
      for (size_t i = 0; i < S.chains.size(); i++) {
        SvChain& x = S.chains[i];
        if (!sv_chain_ready(x, false)) continue;
        x.busy = true;
        id = x.queue.front();
        S.reqs[id].t_start = mono_time();
        r = S.reqs[id];
        c.n = x.n; c.degree = x.degree; c.d_adj = x.d_adj;
        break;
      }
    }
    if (!r.swaps.empty()) {
      if (r.swaps.size() > s_cap) { dfree(&d_swaps); dalloc(&d_swaps, r.swaps.size()); s_cap = r.swaps.size(); }
      SwapJob J;
      J.adj = c.d_adj; J.nbr16 = nullptr; J.n = c.n; J.degree = c.degree; J.dp16 = 0; J.first = 0;
      J.nswaps = (int)(r.swaps.size() / 4);
      HIPC(hipMemcpyAsync(d_swaps, r.swaps.data(), sizeof(int) * r.swaps.size(), hipMemcpyHostToDevice, ev.st));
      HIPC(hipMemcpyAsync(d_job, &J, sizeof(SwapJob), hipMemcpyHostToDevice, ev.st));
      hipLaunchKernelGGL(k_apply_swaps, dim3(1), dim3(64), 0, ev.st, d_job, 1, d_swaps, d_status);
      HIPC(hipGetLastError());
      HIPC(hipMemcpyAsync(h_status, d_status, sizeof(int), hipMemcpyDeviceToHost, ev.st));
      HIPC(hipStreamSynchronize(ev.st));
    } else {
      *h_status = 0;
    }
    if (c.n != shaped_n || c.degree != shaped_d) {
      ev.shape(0, c.n, c.degree, 1, 0, -1, nullptr, 0);
      ODPX_Plan p{};
      p.plan = ODPX_PLAN_LEVEL; p.kw = LEVEL_BLOCK_WORDS; p.upw = 8; p.seg_major = 0; p.xcd = 1; p.group = 0;
      ev.apply_plan(p);
      shaped_n = c.n; shaped_d = c.degree;
    }
    ev.ext_adj = c.d_adj;
    int d; long sm; double as;
    ev.evaluate(&d, &sm, &as);
    if (d > c.n) { d = INT_MAX; sm = LONG_MAX; as = DBL_MAX; }
    double now = mono_time();
    {
      std::lock_guard<std::mutex> lk(S.mu);
      SvReq& q = S.reqs[id];
      q.diameter = d; q.sum = sm; q.aspl = as; q.status = *h_status ? -1 : 0; q.t_done = now;
      sv_complete(id, q);
    }
    S.cv_done.notify_all();
    S.cv_work.notify_all();
  }
  dfree(&d_job); dfree(&d_swaps); dfree(&d_status); hfree(&h_status);
  hipStream_t st = ev.st;
  ev.release();
  HIPC(hipStreamDestroy(st));
}

// bytes of level-plan bit buffers one large worker holds for the largest chain size
static size_t sv_level_bytes(int n)
{
  int W = (n + 63) / 64;
  int kw = std::min(LEVEL_BLOCK_WORDS, W);
  size_t blocks = (W + kw - 1) / kw;
  return blocks * 2 * (size_t)n * kw * 8 + (size_t)n * round_up(64, 4) * 4;
}

extern "C" int ODPX_Serve_start(int small_workers, int large_workers, double hbm_budget_gb)
{
  ctx_init();
  if (S.running) return -1;
  S.stopping = false;
  S.budget = (size_t)(hbm_budget_gb * 1e9);
  S.used = 0;
  S.next_id = 1;
  S.outstanding = 0;
  S.large_slots = std::max(0, large_workers);
  for (int i = 0; i < std::max(1, small_workers); i++) S.workers.emplace_back(sv_small_worker);
  S.running = true;
  return 0;
}

extern "C" int ODPX_Serve_add_chain(int nodes, int degree, const int* adjacency)
{
  if (!S.running || nodes < 2 || degree < 1) return -1;
  SvChain c;
  c.n = nodes; c.degree = degree; c.dp16 = round_up(degree, 8);
  c.wb = batch_wb(nodes);
  c.small = c.wb > 0;
  c.bytes = (size_t)nodes * degree * 4 + (c.small ? (size_t)nodes * c.dp16 * 2 : 0);
  size_t extra = 0;
  std::lock_guard<std::mutex> lk(S.mu);
  if (!c.small) {
    // the first large chain of a size brings the large workers and their buffers
    bool have = false;
    for (const SvChain& x : S.chains) if (!x.small && x.n >= nodes) have = true;
    if (!have && S.large_slots > 0) extra = (size_t)S.large_slots * sv_level_bytes(nodes);
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
  if (extra) {
    int have_workers = 0;
    for (const SvChain& x : S.chains) if (!x.small) have_workers = 1;
    if (!have_workers)
      for (int i = 0; i < S.large_slots; i++) S.workers.emplace_back(sv_large_worker);
  }
  S.chains.push_back(c);
  return (int)S.chains.size() - 1;
}

extern "C" long long ODPX_Serve_submit(int chain, int nswaps, const int* swaps)
{
  long long id;
  {
    std::lock_guard<std::mutex> lk(S.mu);
    if (!S.running || chain < 0 || chain >= (int)S.chains.size() || nswaps < 0) return -1;
    if (!S.chains[chain].small && S.large_slots == 0) return -1;
    id = S.next_id++;
    SvReq r;
    r.chain = chain;
    r.swaps.assign(swaps, swaps + 4 * nswaps);
    r.t_submit = mono_time();
    S.reqs[id] = r;
    S.chains[chain].queue.push_back(id);
    S.outstanding++;
  }
  S.cv_work.notify_all();
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
  std::lock_guard<std::mutex> lk(S.mu);
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
  std::unique_lock<std::mutex> lk(S.mu);
  auto done = [&] {
    for (long long c : S.completed) if (c == id) return true;
    return false;
  };
  if (S.reqs.find(id) == S.reqs.end()) return -1;
  S.cv_done.wait(lk, done);
  for (auto it = S.completed.begin(); it != S.completed.end(); ++it)
    if (*it == id) { S.completed.erase(it); break; }
  auto it = S.reqs.find(id);
  sv_fill(id, it->second, out);
  S.reqs.erase(it);
  return 0;
}

extern "C" void ODPX_Serve_drain(void)
{
  std::unique_lock<std::mutex> lk(S.mu);
  S.cv_done.wait(lk, [] { return S.outstanding == 0; });
}

extern "C" void ODPX_Serve_stop(void)
{
  if (!S.running) return;
  ODPX_Serve_drain();
  {
    std::lock_guard<std::mutex> lk(S.mu);
    S.stopping = true;
  }
  S.cv_work.notify_all();
  for (std::thread& t : S.workers) t.join();
  S.workers.clear();
  for (SvChain& c : S.chains) { dfree(&c.d_adj); dfree(&c.d_nbr16); }
  S.chains.clear();
  S.reqs.clear();
  S.completed.clear();
  S.running = false;
}
