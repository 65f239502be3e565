// Request server of the legalizer (lg_srv_* in include/skyline_lg.h): a
// continuous-batching loop on one GPU.  No counterpart in the reference,
// whose legalizer handles one design per process (Legalizer::run,
// Legalizer.cpp:197-228).  Included at the end of lg_device.hip (same
// translation unit as the kernels and the handle code).
//
// Loop of the server thread:
//   1. admit queued requests in arrival order while a slot is free and the
//      request's device buffers fit the budget (pool of device memory);
//      upload them and run their solve set-up (rho, lambda0, duals, flags);
//   2. when the set of resident requests changed, rebuild the slot table
//      (one Args per slot) and the block maps of the batched launches
//      (256-cell blocks, one block per slot, per partition, pixel blocks);
//   3. launch `iters_per_sync` ADMM iterations, each as one batched launch
//      per step covering every resident request (kb_* kernels);
//   4. copy every slot's Ctl to the host in one transfer, synchronise, and
//      retire the slots whose Ctl says done (final refine, rotation, QP,
//      checkPlace, positions to the host), freeing their slot and memory.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

namespace {

using SrvClock = std::chrono::steady_clock;

struct SrvRequest {
  int64_t ticket = 0;
  HostPrep prep;
  size_t bytes = 0;
  int state = LG_SRV_QUEUED;
  int status = LG_OK;
  std::string err;
  SrvClock::time_point t_submit, t_admit, t_done;
  std::vector<int32_t> lx, ly, rot;
  LgStats stats{};
  int num_cells = 0;
  bool waited = false;  // a thread is inside lg_srv_wait for this ticket
};

struct SrvSlot {
  SrvRequest* req = nullptr;
  LgHandle* h = nullptr;
};

// A batched launch's block range in the map buffer.
struct MapRange {
  int off = 0, n = 0;
};

double secs(SrvClock::time_point a, SrvClock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

}  // namespace

struct LgServer {
  LgSrvConfig cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // the batched dual pass, beside the partition kernels
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // one batched iteration as a HIP graph (cfg.use_graphs), captured on `cap`
  // whenever the tables change
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_stale = true;
  DevPool pool;
  Ctl* d_ctl = nullptr;   // [max_slots] iteration control of every slot
  Ctl* h_ctl = nullptr;   // pinned host copy
  Args* d_tab = nullptr;  // [max_slots] Args of every slot
  Args* h_tab = nullptr;  // pinned staging
  int2* d_map = nullptr;  // block maps of the batched launches
  int2* h_map = nullptr;  // pinned staging
  size_t map_cap = 0;
  std::vector<SrvSlot> slots;
  std::vector<int> active;  // resident slots, in admission order
  bool dirty = true;
  // geometry of the current tables
  MapRange m256, m1, mk, mpix, mbk;
  int any_adaptive = 0, any_reg = 0, any_nonreg = 0, top_slot = 0;
  size_t lds_reg = 0, lds_small = 0, lds_big = 0;
  // queue and request registry (guarded by mu)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvRequest*> queue;
  std::unordered_map<int64_t, SrvRequest*> reqs;
  int waiters = 0;  // threads inside lg_srv_wait (lg_srv_destroy waits for them)
  int64_t next_ticket = 1;
  bool stop = false;
  LgSrvCounters ctr{};
  std::string fatal;  // a device error of the loop (every later request fails)
  std::thread worker;
  // primal-step timing (lg_srv_profile): two events per batched iteration
  // bracket its partition-kernel launches; bytes are the algorithmic bytes
  // (12 per pixel + 24 per cell) of the iterations each slot executed
  int prof = 0;        // set by lg_srv_profile (under mu)
  int prof_batch = 0;  // its value at the start of the current batch
  std::vector<hipEvent_t> ev;
  std::vector<int32_t> iter_before;
  double prof_ms = 0, prof_bytes = 0;
  int64_t prof_launches = 0;
};

namespace {

int srv_grow_map(LgServer* S, size_t need) {
  if (need <= S->map_cap) return LG_OK;
  size_t cap = std::max(need, S->map_cap * 2);
  if (S->d_map) (void)hipFree(S->d_map);
  if (S->h_map) (void)hipHostFree(S->h_map);
  S->d_map = nullptr;
  S->h_map = nullptr;
  S->map_cap = 0;
  HIP_TRY(hipMalloc((void**)&S->d_map, cap * sizeof(int2)));
  HIP_TRY(hipHostMalloc((void**)&S->h_map, cap * sizeof(int2), hipHostMallocDefault));
  S->map_cap = cap;
  return LG_OK;
}

// Slot table and block maps of the resident requests (server stream idle).
int srv_build_tables(LgServer* S) {
  std::vector<int2> m[5];
  S->any_adaptive = S->any_reg = S->any_nonreg = 0;
  S->lds_reg = S->lds_small = S->lds_big = 0;
  S->top_slot = 0;
  for (int sl : S->active) {
    LgHandle* h = S->slots[sl].h;
    const Args& a = h->a;
    S->h_tab[sl] = a;
    S->top_slot = std::max(S->top_slot, sl + 1);
    for (int b = 0; b < (a.N + 255) / 256; b++) m[0].push_back(make_int2(sl, b));
    m[1].push_back(make_int2(sl, 0));
    for (int b = 0; b < a.K; b++) m[2].push_back(make_int2(sl, b));
    const int gx = (a.x_size + 1023) / 1024;
    for (int b = 0; b < gx * a.y_size; b++) m[3].push_back(make_int2(sl, b));
    for (int b = 0; b < (a.N + kBucketCells - 1) / kBucketCells; b++) m[4].push_back(make_int2(sl, b));
    S->any_adaptive |= a.use_adaptive ? 1 : 0;
    const int reg = h->p.sorting_algorithm == LG_SORT_M_BITONIC || h->p.sorting_algorithm == LG_SORT_S_BITONIC;
    if (reg) {
      S->any_reg = 1;
      S->lds_reg = std::max(S->lds_reg, partition_lds_bytes(h, kNpadSmall, 0));
    } else {
      S->any_nonreg = 1;
      S->lds_small = std::max(S->lds_small, partition_lds_bytes(h, kNpadSmall, 1));
    }
    S->lds_big = std::max(S->lds_big, partition_lds_bytes(h, kMaxNpadLds, 1));
  }
  size_t total = 0;
  MapRange* r[5] = {&S->m256, &S->m1, &S->mk, &S->mpix, &S->mbk};
  for (int i = 0; i < 5; i++) {
    r[i]->off = (int)total;
    r[i]->n = (int)m[i].size();
    total += m[i].size();
  }
  if (total > (size_t)INT_MAX) return fail(LG_ERR_MEMORY, "request server: block maps too large");
  int st = srv_grow_map(S, total);
  if (st) return st;
  for (int i = 0; i < 5; i++)
    if (!m[i].empty()) std::memcpy(S->h_map + r[i]->off, m[i].data(), m[i].size() * sizeof(int2));
  HIP_TRY(hipMemcpyAsync(S->d_tab, S->h_tab, S->top_slot * sizeof(Args), hipMemcpyHostToDevice, S->stream));
  HIP_TRY(hipMemcpyAsync(S->d_map, S->h_map, total * sizeof(int2), hipMemcpyHostToDevice, S->stream));
  S->dirty = false;
  // the launch geometry changed: the graph of an iteration is captured anew
  S->graph_stale = true;
  return LG_OK;
}

// One ADMM iteration of every resident request (the steps of
// launch_iteration, each as one batched launch).  The dual pass (with lazy
// duals the catch-up pass, followed by the stamps) runs when `pass` is set.
int srv_launch_iteration(LgServer* S, int it, hipStream_t s, bool pass) {
  const Args* tab = S->d_tab;
  const int2* mp = S->d_map;
  auto at = [&](const MapRange& r) { return mp + r.off; };
  if (S->mbk.n > 0) hipLaunchKernelGGL(kb_bucket, dim3(S->mbk.n), dim3(256), 0, s, tab, at(S->mbk));
  hipLaunchKernelGGL(kb_scan, dim3(S->m1.n), dim3(1024), 0, s, tab, at(S->m1));
  // the dual pass on the side stream, beside the partition kernels
  if (pass) {
    HIP_TRY(hipEventRecord(S->ev_fork, s));
    HIP_TRY(hipStreamWaitEvent(S->side, S->ev_fork, 0));
    hipLaunchKernelGGL(kb_dual_pass, dim3(S->mpix.n), dim3(256), 0, S->side, tab, at(S->mpix));
    hipLaunchKernelGGL(kb_stamp_inactive, dim3(S->m1.n), dim3(256), 0, S->side, tab, at(S->m1));
    HIP_TRY(hipEventRecord(S->ev_join, S->side));
  }
  if (S->any_adaptive) {
    if (S->m256.n > 0)
      hipLaunchKernelGGL(kb_extract_congested, dim3(S->m256.n), dim3(256), 0, s, tab, at(S->m256));
    hipLaunchKernelGGL(kb_adaptive, dim3(S->m1.n), dim3(1024), 0, s, tab, at(S->m1));
  }
  hipEvent_t* e = nullptr;
  if (S->prof_batch && s == S->stream) {
    e = &S->ev[2 * it];
    HIP_TRY(hipEventRecord(e[0], s));
  }
  if (S->any_reg)
    hipLaunchKernelGGL((kb_partition<MODE_PRIMAL, 0, kPartWaves>), dim3(S->mk.n), dim3(kPartWaves * kWave), S->lds_reg,
                       s, tab, at(S->mk), 0,
                       kNpadSmall, 1);
  if (S->any_nonreg)
    hipLaunchKernelGGL((kb_partition<MODE_PRIMAL, 1, 1>), dim3(S->mk.n), dim3(kWave), S->lds_small, s, tab, at(S->mk),
                       0, kNpadSmall, 0);
  // the large class also takes the partitions the four-wave launch deferred
  hipLaunchKernelGGL((kb_partition<MODE_PRIMAL, 1, 1>), dim3(S->mk.n), dim3(kWave), S->lds_big, s, tab, at(S->mk), 1,
                     kMaxNpad, -1);
  if (e) HIP_TRY(hipEventRecord(e[1], s));
  if (pass) HIP_TRY(hipStreamWaitEvent(s, S->ev_join, 0));
  hipLaunchKernelGGL(kb_finish, dim3(S->m1.n), dim3(1), 0, s, tab, at(S->m1));
  HIP_TRY(hipGetLastError());
  return LG_OK;
}

// One batched iteration captured into a graph (tables built, nothing runs)
// and instantiated.
int srv_capture(LgServer* S) {
  HIP_TRY(hipStreamBeginCapture(S->cap, hipStreamCaptureModeThreadLocal));
  // every captured iteration carries the dual / catch-up pass
  const int st = srv_launch_iteration(S, 0, S->cap, true);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(S->cap, &g);
  if (st || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st ? st : fail(LG_ERR_DEVICE, std::string("request server: capture: ") + hipGetErrorString(e));
  }
  if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
  S->gexec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&S->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    S->gexec = nullptr;
    return fail(LG_ERR_DEVICE, std::string("request server: hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  S->graph_stale = false;
  std::lock_guard<std::mutex> lk(S->mu);
  S->ctr.graph_captures++;
  return LG_OK;
}

void srv_complete(LgServer* S, SrvRequest* r, int status, const std::string& err) {
  std::lock_guard<std::mutex> lk(S->mu);
  r->status = status;
  r->err = err;
  r->state = status == LG_OK ? LG_SRV_DONE : LG_SRV_FAILED;
  r->t_done = SrvClock::now();
  if (status == LG_OK) S->ctr.completed++;
  else S->ctr.failed++;
  S->cv_done.notify_all();
}

// Final refine, post-process, checkPlace and copy-out of a finished handle;
// then the handle's memory goes back to the pool.
int srv_finish_handle(LgServer* S, SrvRequest* r, LgHandle* h) {
  int st = solve_end(h, S->stream);
  if (!st) st = lg_postprocess(h, S->stream);
  if (!st) st = lg_check(h, nullptr);
  if (!st) {
    const int n = h->a.N;
    r->lx.resize(n);
    r->ly.resize(n);
    r->rot.resize(n);
    st = lg_copy_out(h, r->lx.data(), r->ly.data(), r->rot.data());
  }
  if (!st) st = lg_get_stats(h, &r->stats);
  return st;
}

void srv_retire(LgServer* S, int sl) {
  SrvSlot& slot = S->slots[sl];
  LgHandle* h = slot.h;
  SrvRequest* r = slot.req;
  h->host_ctl = S->h_ctl[sl];
  int st = srv_finish_handle(S, r, h);
  const std::string err = st ? g_err : std::string();
  lg_destroy(h);
  slot = SrvSlot{};
  S->active.erase(std::find(S->active.begin(), S->active.end(), sl));
  S->dirty = true;
  srv_complete(S, r, st, err);
}

// Creates the request's handle in slot `sl` and starts its solve; returns
// LG_ERR_MEMORY (request untouched) when the pool cannot hold it now.
int srv_admit(LgServer* S, SrvRequest* r, int sl) {
  LgHandle* h = nullptr;
  int st = create_from_prep(r->prep, &S->pool, &S->d_ctl[sl], S->stream, &h);
  if (st) return st;
  r->t_admit = SrvClock::now();
  r->num_cells = r->prep.N;
  r->prep = HostPrep();  // host copies no longer needed
  {
    std::lock_guard<std::mutex> lk(S->mu);
    r->state = LG_SRV_RUNNING;
  }
  if (h->a.N == 0 || h->p.max_admm_iter <= 0) {
    // nothing to iterate: the solo path, then retire at once
    st = lg_solve(h, S->stream);
    if (!st) st = lg_postprocess(h, S->stream);
    if (!st) st = lg_check(h, nullptr);
    if (!st) {
      r->lx.assign(h->a.N, 0);
      r->ly.assign(h->a.N, 0);
      r->rot.assign(h->a.N, 0);
      st = lg_copy_out(h, r->lx.data(), r->ly.data(), r->rot.data());
    }
    if (!st) st = lg_get_stats(h, &r->stats);
    const std::string err = st ? g_err : std::string();
    lg_destroy(h);
    srv_complete(S, r, st, err);
    return LG_OK;
  }
  if ((st = solve_begin(h, S->stream))) {
    const std::string err = g_err;
    lg_destroy(h);
    srv_complete(S, r, st, err);
    return LG_OK;
  }
  S->slots[sl] = SrvSlot{r, h};
  S->h_ctl[sl].iter = 0;
  S->active.push_back(sl);
  S->dirty = true;
  return LG_OK;
}

void srv_fail_all(LgServer* S, const std::string& why) {
  (void)hipStreamSynchronize(S->stream);
  for (int sl : std::vector<int>(S->active)) {
    SrvSlot& slot = S->slots[sl];
    lg_destroy(slot.h);
    srv_complete(S, slot.req, LG_ERR_STATE, why);
    slot = SrvSlot{};
  }
  S->active.clear();
  S->dirty = true;
  std::deque<SrvRequest*> q;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    q.swap(S->queue);
  }
  for (SrvRequest* r : q) srv_complete(S, r, LG_ERR_STATE, why);
}

void srv_loop(LgServer* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    std::vector<SrvRequest*> take;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->active.empty(); });
      if (S->stop) break;
      int free_slots = S->cfg.max_slots - (int)S->active.size();
      size_t avail = S->pool.available();
      while (!S->queue.empty() && free_slots > 0 && S->queue.front()->bytes <= avail) {
        take.push_back(S->queue.front());
        avail -= S->queue.front()->bytes;
        S->queue.pop_front();
        free_slots--;
      }
    }
    for (size_t i = 0; i < take.size(); i++) {
      int sl = 0;
      while (S->slots[sl].h) sl++;
      SrvRequest* r = take[i];
      int st = srv_admit(S, r, sl);
      if (st == LG_ERR_MEMORY && S->active.empty() && S->pool.live.empty()) {
        // nothing resident: chunks too small for this request are given
        // back and the pool starts again
        (void)hipStreamSynchronize(S->stream);
        S->pool.destroy();
        st = srv_admit(S, r, sl);
      }
      if (st == LG_ERR_MEMORY && !S->active.empty()) {
        // fragmentation: wait for a retirement, keep arrival order
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t j = take.size(); j > i; j--) S->queue.push_front(take[j - 1]);
        break;
      }
      if (st) srv_complete(S, r, st, g_err);
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->ctr.peak_slots = std::max<int32_t>(S->ctr.peak_slots, (int32_t)S->active.size());
      S->prof_batch = S->prof;
    }
    if (S->active.empty()) continue;
    int st = S->dirty ? srv_build_tables(S) : LG_OK;
    if (!st && S->prof_batch) {
      // iteration counters at the start of the batch
      S->iter_before.assign(S->top_slot, 0);
      for (int sl : S->active) S->iter_before[sl] = S->h_ctl[sl].iter;
    }
    if (!st && S->cfg.use_graphs && !S->prof_batch) {
      if (!S->gexec || S->graph_stale) st = srv_capture(S);
      for (int it = 0; !st && it < S->cfg.iters_per_sync; it++)
        if (hipGraphLaunch(S->gexec, S->stream) != hipSuccess) st = fail(LG_ERR_DEVICE, "request server: graph launch");
    } else {
      for (int it = 0; !st && it < S->cfg.iters_per_sync; it++) {
        // lazy duals: the catch-up pass every kCatchupPeriod batched
        // iterations (every iteration while a slot runs the adaptive search,
        // whose prices must stay current)
        const int64_t n = S->ctr.iterations + it;
        const bool pass = S->any_adaptive || n % kCatchupPeriod == kCatchupPeriod - 1;
        st = srv_launch_iteration(S, it, S->stream, pass);
      }
    }
    if (!st && hipMemcpyAsync(S->h_ctl, S->d_ctl, S->top_slot * sizeof(Ctl), hipMemcpyDeviceToHost, S->stream) !=
                   hipSuccess)
      st = fail(LG_ERR_DEVICE, "request server: control copy failed");
    if (!st && hipStreamSynchronize(S->stream) != hipSuccess)
      st = fail(LG_ERR_DEVICE, "request server: device error in the batched loop");
    if (st) {
      S->fatal = g_err;
      srv_fail_all(S, S->fatal);
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
      break;
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->ctr.batches++;
      S->ctr.iterations += S->cfg.iters_per_sync;
      if (S->prof_batch) {
        for (int it = 0; it < S->cfg.iters_per_sync; it++) {
          float ms = 0;
          if (hipEventElapsedTime(&ms, S->ev[2 * it], S->ev[2 * it + 1]) == hipSuccess) S->prof_ms += ms;
        }
        S->prof_launches += S->cfg.iters_per_sync;
        for (int sl : S->active) {
          const Args& a = S->slots[sl].h->a;
          const double ran = S->h_ctl[sl].iter - S->iter_before[sl];
          S->prof_bytes += ran * (12.0 * a.x_size * a.y_size + 24.0 * a.N);
        }
      }
    }
    for (int sl : std::vector<int>(S->active))
      if (S->h_ctl[sl].done || S->h_ctl[sl].error) srv_retire(S, sl);
  }
  srv_fail_all(S, S->fatal.empty() ? "request server stopped" : S->fatal);
}

}  // namespace

extern "C" {

void lg_srv_default_config(LgSrvConfig* c) {
  if (!c) return;
  c->max_slots = 64;
  c->iters_per_sync = 16;
  c->mem_budget_bytes = 200ll * 1000 * 1000 * 1000;
  c->pool_chunk_bytes = 256ll << 20;
  c->use_graphs = 0;
  c->reserved = 0;
}

int lg_srv_create(const LgSrvConfig* c, LgServer** out) {
  if (!c || !out) return fail(LG_ERR_ARG, "null argument");
  if (c->max_slots < 1 || c->max_slots > 1024 || c->iters_per_sync < 1 || c->iters_per_sync > 1024 ||
      c->mem_budget_bytes <= 0 || c->pool_chunk_bytes <= 0)
    return fail(LG_ERR_ARG, "bad server configuration");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LG_ERR_DEVICE, "no HIP device");
  int st = upload_directions();
  if (st) return st;
  LgServer* S = new LgServer();
  S->cfg = *c;
  auto bail = [&](int code) {
    std::string msg = g_err;
    if (S->d_ctl) (void)hipFree(S->d_ctl);
    if (S->d_tab) (void)hipFree(S->d_tab);
    if (S->h_ctl) (void)hipHostFree(S->h_ctl);
    if (S->h_tab) (void)hipHostFree(S->h_tab);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    if (S->side) (void)hipStreamDestroy(S->side);
    if (S->cap) (void)hipStreamDestroy(S->cap);
    if (S->ev_fork) (void)hipEventDestroy(S->ev_fork);
    if (S->ev_join) (void)hipEventDestroy(S->ev_join);
    delete S;
    g_err = msg;
    return code;
  };
  if (hipGetDevice(&S->device) != hipSuccess) return bail(fail(LG_ERR_DEVICE, "hipGetDevice"));
  S->pool.budget = (size_t)c->mem_budget_bytes;
  S->pool.chunk_bytes = (size_t)c->pool_chunk_bytes;
  const size_t ns = (size_t)c->max_slots;
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&S->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&S->cap, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&S->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&S->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void**)&S->d_ctl, ns * sizeof(Ctl)) != hipSuccess ||
      hipMalloc((void**)&S->d_tab, ns * sizeof(Args)) != hipSuccess ||
      hipHostMalloc((void**)&S->h_ctl, ns * sizeof(Ctl), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&S->h_tab, ns * sizeof(Args), hipHostMallocDefault) != hipSuccess)
    return bail(fail(LG_ERR_MEMORY, "request server: allocation of the slot tables failed"));
  std::memset(S->h_ctl, 0, ns * sizeof(Ctl));
  std::memset(S->h_tab, 0, ns * sizeof(Args));
  S->slots.resize(ns);
  S->worker = std::thread(srv_loop, S);
  *out = S;
  return LG_OK;
}

int lg_srv_submit(LgServer* S, const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, int64_t* ticket) {
  if (!S || !ticket) return fail(LG_ERR_ARG, "null argument");
  SrvRequest* r = new SrvRequest();
  r->t_submit = SrvClock::now();
  int st = prepare_host(g, c, p, r->prep);
  if (st) {
    delete r;
    return st;
  }
  if (r->prep.tech) {
    delete r;
    return fail(LG_ERR_ARG, "request server: pin / edge-spacing technology is legalized by lg_create .. lg_solve only");
  }
  r->bytes = handle_bytes(r->prep);
  r->num_cells = r->prep.N;
  if (r->bytes > S->pool.budget) {
    delete r;
    return fail(LG_ERR_MEMORY, "request larger than the server's device-memory budget");
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      delete r;
      return fail(LG_ERR_STATE, S->fatal.empty() ? "request server stopped" : S->fatal);
    }
    r->ticket = S->next_ticket++;
    S->reqs.emplace(r->ticket, r);
    S->queue.push_back(r);
    S->ctr.submitted++;
    *ticket = r->ticket;
  }
  S->cv_work.notify_one();
  return LG_OK;
}

int lg_srv_poll(LgServer* S, int64_t ticket, int32_t* state) {
  if (!S || !state) return fail(LG_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return fail(LG_ERR_ARG, "unknown ticket");
  *state = it->second->state;
  return LG_OK;
}

int lg_srv_wait(LgServer* S, int64_t ticket, LgSrvResult* res) {
  if (!S || !res) return fail(LG_ERR_ARG, "null argument");
  SrvRequest* r = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(LG_ERR_ARG, "unknown ticket");
    r = it->second;
    if (r->waited) return fail(LG_ERR_STATE, "another thread is waiting for this ticket");
    r->waited = true;
    S->waiters++;
    S->cv_done.wait(lk, [&] { return r->state == LG_SRV_DONE || r->state == LG_SRV_FAILED; });
    // by key: a submit during the wait may have rehashed the map
    S->reqs.erase(ticket);
    // the record is this thread's from here on; the server is not touched
    // again after the lock is released (lg_srv_destroy may free it then)
    S->waiters--;
    S->cv_done.notify_all();
  }
  res->num_cells = r->num_cells;
  res->status = r->status;
  res->stats = r->stats;
  const bool admitted = r->t_admit.time_since_epoch().count() != 0;
  res->queued_s = admitted ? secs(r->t_submit, r->t_admit) : secs(r->t_submit, r->t_done);
  res->run_s = admitted ? secs(r->t_admit, r->t_done) : 0.0;
  res->latency_s = secs(r->t_submit, r->t_done);
  res->stats.t_admm_s = res->run_s;
  res->stats.t_total_s = res->latency_s;
  if (r->status == LG_OK) {
    const size_t n = r->lx.size() * sizeof(int32_t);
    if (res->out_lx && n) std::memcpy(res->out_lx, r->lx.data(), n);
    if (res->out_ly && n) std::memcpy(res->out_ly, r->ly.data(), n);
    if (res->out_rot && n) std::memcpy(res->out_rot, r->rot.data(), n);
  }
  const int st = r->status;
  const std::string err = r->err;
  delete r;
  if (st) return fail(st, err);
  return LG_OK;
}

int lg_srv_counters(LgServer* S, LgSrvCounters* out) {
  if (!S || !out) return fail(LG_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->ctr;
  out->active = (int32_t)S->active.size();
  out->queued = (int32_t)S->queue.size();
  out->pool_reserved_bytes = (int64_t)S->pool.reserved;
  out->pool_peak_bytes = (int64_t)S->pool.peak;
  return LG_OK;
}

int lg_srv_profile(LgServer* S, int enable) {
  if (!S) return fail(LG_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  if (enable && S->ev.empty()) {
    S->ev.resize(2 * (size_t)S->cfg.iters_per_sync);
    for (auto& e : S->ev) HIP_TRY(hipEventCreate(&e));
  }
  S->prof = enable ? 1 : 0;
  S->prof_ms = S->prof_bytes = 0;
  S->prof_launches = 0;
  return LG_OK;
}

int lg_srv_profile_read(LgServer* S, double* ms, int64_t* iterations, double* bytes) {
  if (!S) return fail(LG_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  if (ms) *ms = S->prof_ms;
  if (iterations) *iterations = S->prof_launches;
  if (bytes) *bytes = S->prof_bytes;
  return LG_OK;
}

int lg_srv_destroy(LgServer* S) {
  if (!S) return LG_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  {
    // every request is done or failed now (the loop's srv_fail_all); wait
    // until the threads blocked in lg_srv_wait have taken their records,
    // then drop the records nobody collected
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.notify_all();
    S->cv_done.wait(lk, [&] { return S->waiters == 0; });
    for (auto& kv : S->reqs) delete kv.second;
    S->reqs.clear();
  }
  S->pool.destroy();
  if (S->d_map) (void)hipFree(S->d_map);
  if (S->h_map) (void)hipHostFree(S->h_map);
  (void)hipFree(S->d_ctl);
  (void)hipFree(S->d_tab);
  (void)hipHostFree(S->h_ctl);
  (void)hipHostFree(S->h_tab);
  for (auto& e : S->ev) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(S->side);
  (void)hipStreamDestroy(S->side);
  if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
  (void)hipStreamDestroy(S->cap);
  (void)hipEventDestroy(S->ev_fork);
  (void)hipEventDestroy(S->ev_join);
  (void)hipStreamDestroy(S->stream);
  delete S;
  return LG_OK;
}

}  // extern "C"
