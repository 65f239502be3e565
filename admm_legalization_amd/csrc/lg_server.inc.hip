// Request server of the legalizer (lg_srv_* in include/skyline_lg.h): a
// continuous-batching loop on one GPU.  No counterpart in the reference,
// whose legalizer handles one design per process (Legalizer::run,
// Legalizer.cpp:197-228).  Included at the end of lg_device.hip (same
// translation unit as the kernels and the handle code).
//
// Resident requests iterate in lanes.  With LgSrvConfig::short_cells > 0
// there are two: requests with fewer cells than that iterate in the short
// lane, the others in the long lane.  Each lane has its own streams, slots,
// slot table, block maps and batch cadence, so a short request advances at
// the pace of the short lane's launches, not at that of the long requests'
// partition launches; the two lanes' kernels run side by side on the GPU.
// With short_cells == 0 there is one lane.
//
// Loop of the server thread:
//   1. admit queued requests while fewer than max_slots are resident, the
//      request's device buffers fit the budget (pool of device memory) and
//      its lane is between batches: in arrival order, or with
//      max_overtakes > 0 the smallest first, except that a request passed
//      over by max_overtakes later requests goes before any other; upload
//      them and run their solve set-up (rho, lambda0, duals, flags);
//   2. for every lane between batches with resident requests: when its set
//      of requests changed, rebuild its slot table (one Args per slot) and
//      block maps (256-cell blocks, one block per slot, per partition, pixel
//      blocks); enqueue `iters_per_sync` ADMM iterations, each as one batched
//      launch per step covering every request of the lane (kb_* kernels),
//      the copy of the lane's Ctls to the host and an event;
//   3. wait until a lane's event completes (or a request arrives while a
//      lane is between batches); retire that lane's slots whose Ctl says
//      done (final refine, rotation, QP, checkPlace, positions to the host),
//      freeing their slot and memory.
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
  int lane = 0;
  int passed = 0;       // later requests admitted while this one was queued
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

// Resident requests iterated together on their own streams.
struct SrvLane {
  int slot_lo = 0;  // the lane's slots are [slot_lo, slot_lo + max_slots)
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // the batched dual pass, beside the partition kernels
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_done = nullptr;  // end of the batch in flight (Ctls on the host)
  // one batched iteration as a HIP graph (use_graphs), captured whenever
  // the lane's tables change
  hipGraphExec_t gexec = nullptr;
  bool graph_stale = true;
  int2* d_map = nullptr;  // block maps of the batched launches
  int2* h_map = nullptr;  // pinned staging
  size_t map_cap = 0;
  std::vector<int> active;  // resident slots, in admission order
  bool dirty = true;
  bool in_flight = false;
  // geometry of the current tables (top_slot: one past the highest slot)
  MapRange m256, m1, mk, mpix, mbk, mlg;
  int any_adaptive = 0, any_reg = 0, any_nonreg = 0, top_slot = 0;
  int large_k = 0;  // a request with more than 8192 partitions (the scan's large path)
  int max_k = 0;    // most partitions of a resident request
  size_t lds_reg = 0, lds_small = 0, lds_big = 0;
  int64_t iterations = 0;  // batched iterations launched in this lane
  // primal-step timing (lg_srv_profile): two events per batched iteration
  // bracket its partition-kernel launches
  int prof_batch = 0;  // the server's `prof` when the batch was launched
  std::vector<hipEvent_t> ev;
  std::vector<int32_t> iter_before;
};

double secs(SrvClock::time_point a, SrvClock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

}  // namespace

struct LgServer {
  LgSrvConfig cfg{};
  int device = 0;
  hipStream_t cap = nullptr;  // stream the iteration graphs are captured on
  DevPool pool;
  std::vector<SrvLane> lanes;
  Ctl* d_ctl = nullptr;   // [lanes * max_slots] iteration control of every slot
  Ctl* h_ctl = nullptr;   // pinned host copy
  Args* d_tab = nullptr;  // [lanes * max_slots] Args of every slot
  Args* h_tab = nullptr;  // pinned staging
  std::vector<SrvSlot> slots;
  int resident = 0;  // requests in slots, over all lanes
  // queue (arrival order) and request registry (guarded by mu)
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
  // lg_srv_profile: summed partition-launch time, bracketed iterations, the
  // algorithmic bytes (12 per pixel + 24 per cell) of the iterations each
  // slot executed in them, and every launch interval in milliseconds from
  // prof_ref (recorded on an idle stream when profiling starts), for the
  // time during which any lane's partition launches ran
  int prof = 0;  // set by lg_srv_profile (under mu)
  double prof_ms = 0, prof_bytes = 0;
  int64_t prof_launches = 0;
  hipStream_t prof_stream = nullptr;
  hipEvent_t prof_ref = nullptr;
  std::vector<std::pair<double, double>> prof_iv;
};

namespace {

int srv_grow_map(SrvLane& L, size_t need) {
  if (need <= L.map_cap) return LG_OK;
  size_t cap = std::max(need, L.map_cap * 2);
  if (L.d_map) (void)hipFree(L.d_map);
  if (L.h_map) (void)hipHostFree(L.h_map);
  L.d_map = nullptr;
  L.h_map = nullptr;
  L.map_cap = 0;
  HIP_TRY(hipMalloc((void**)&L.d_map, cap * sizeof(int2)));
  HIP_TRY(hipHostMalloc((void**)&L.h_map, cap * sizeof(int2), hipHostMallocDefault));
  L.map_cap = cap;
  return LG_OK;
}

// Slot table and block maps of the lane's requests (lane between batches).
int srv_build_tables(LgServer* S, SrvLane& L) {
  std::vector<int2> m[6];
  L.any_adaptive = L.any_reg = L.any_nonreg = L.large_k = L.max_k = 0;
  L.lds_reg = L.lds_small = L.lds_big = 0;
  L.top_slot = L.slot_lo;
  for (int sl : L.active) {
    LgHandle* h = S->slots[sl].h;
    const Args& a = h->a;
    S->h_tab[sl] = a;
    L.top_slot = std::max(L.top_slot, sl + 1);
    for (int b = 0; b < (a.N + 255) / 256; b++) m[0].push_back(make_int2(sl, b));
    m[1].push_back(make_int2(sl, 0));
    for (int b = 0; b < a.K; b++) m[2].push_back(make_int2(sl, b));
    const int gx = (a.x_size + 1023) / 1024;
    for (int b = 0; b < gx * a.y_size; b++) m[3].push_back(make_int2(sl, b));
    for (int b = 0; b < (a.N + kBucketCells - 1) / kBucketCells; b++) m[4].push_back(make_int2(sl, b));
    for (int b = 0; b < std::min(a.K, kLargeGroups); b++) m[5].push_back(make_int2(sl, b));
    L.any_adaptive |= a.use_adaptive ? 1 : 0;
    L.large_k |= a.K > 8 * 1024 ? 1 : 0;
    L.max_k = std::max(L.max_k, a.K);
    const int reg = h->p.sorting_algorithm == LG_SORT_M_BITONIC || h->p.sorting_algorithm == LG_SORT_S_BITONIC;
    if (reg) {
      L.any_reg = 1;
      L.lds_reg = std::max(L.lds_reg, partition_lds_bytes(h, kNpadSmall, 0));
    } else {
      L.any_nonreg = 1;
      L.lds_small = std::max(L.lds_small, partition_lds_bytes(h, kNpadSmall, 1));
    }
    L.lds_big = std::max(L.lds_big, partition_lds_bytes(h, kMaxNpadLds, 1));
  }
  size_t total = 0;
  MapRange* r[6] = {&L.m256, &L.m1, &L.mk, &L.mpix, &L.mbk, &L.mlg};
  for (int i = 0; i < 6; i++) {
    r[i]->off = (int)total;
    r[i]->n = (int)m[i].size();
    total += m[i].size();
  }
  if (total > (size_t)INT_MAX) return fail(LG_ERR_MEMORY, "request server: block maps too large");
  int st = srv_grow_map(L, total);
  if (st) return st;
  for (int i = 0; i < 6; i++)
    if (!m[i].empty()) std::memcpy(L.h_map + r[i]->off, m[i].data(), m[i].size() * sizeof(int2));
  // only the lane's own entries: the other lane's kernels read theirs meanwhile
  HIP_TRY(hipMemcpyAsync(S->d_tab + L.slot_lo, S->h_tab + L.slot_lo, (L.top_slot - L.slot_lo) * sizeof(Args),
                         hipMemcpyHostToDevice, L.stream));
  HIP_TRY(hipMemcpyAsync(L.d_map, L.h_map, total * sizeof(int2), hipMemcpyHostToDevice, L.stream));
  L.dirty = false;
  // the launch geometry changed: the graph of an iteration is captured anew
  L.graph_stale = true;
  return LG_OK;
}

// One ADMM iteration of every request of the lane (the steps of
// launch_iteration, each as one batched launch).  The dual pass (with lazy
// duals the catch-up pass, followed by the stamps) runs when `pass` is set.
int srv_launch_iteration(LgServer* S, SrvLane& L, int it, hipStream_t s, bool pass) {
  const Args* tab = S->d_tab;
  const int2* mp = L.d_map;
  auto at = [&](const MapRange& r) { return mp + r.off; };
  if (L.mbk.n > 0) hipLaunchKernelGGL(kb_bucket, dim3(L.mbk.n), dim3(256), 0, s, tab, at(L.mbk));
  // the scan's workgroup: 256 threads for up to 2048 partitions, else 1024
  if (L.large_k) hipLaunchKernelGGL((kb_scan<true, 1024>), dim3(L.m1.n), dim3(1024), 0, s, tab, at(L.m1));
  else if (L.max_k > 8 * 256) hipLaunchKernelGGL((kb_scan<false, 1024>), dim3(L.m1.n), dim3(1024), 0, s, tab, at(L.m1));
  else hipLaunchKernelGGL((kb_scan<false, 256>), dim3(L.m1.n), dim3(256), 0, s, tab, at(L.m1));
  // the dual pass on the side stream, beside the partition kernels
  if (pass) {
    HIP_TRY(hipEventRecord(L.ev_fork, s));
    HIP_TRY(hipStreamWaitEvent(L.side, L.ev_fork, 0));
    hipLaunchKernelGGL(kb_dual_pass, dim3(L.mpix.n), dim3(256), 0, L.side, tab, at(L.mpix));
    hipLaunchKernelGGL(kb_stamp_inactive, dim3(L.m1.n), dim3(256), 0, L.side, tab, at(L.m1));
    HIP_TRY(hipEventRecord(L.ev_join, L.side));
  }
  if (L.any_adaptive) {
    if (L.m256.n > 0)
      hipLaunchKernelGGL(kb_extract_congested, dim3(L.m256.n), dim3(256), 0, s, tab, at(L.m256));
    hipLaunchKernelGGL(kb_adaptive, dim3(L.m1.n), dim3(1024), 0, s, tab, at(L.m1));
  }
  hipEvent_t* e = nullptr;
  if (L.prof_batch && s == L.stream) {
    e = &L.ev[2 * it];
    HIP_TRY(hipEventRecord(e[0], s));
  }
  if (L.any_reg)
    hipLaunchKernelGGL((kb_partition<MODE_PRIMAL, 0, kPartWaves>), dim3(L.mk.n), dim3(kPartWaves * kWave), L.lds_reg,
                       s, tab, at(L.mk), 0, kNpadSmall, 1);
  if (L.any_nonreg)
    hipLaunchKernelGGL((kb_partition<MODE_PRIMAL, 1, 1>), dim3(L.mk.n), dim3(kWave), L.lds_small, s, tab, at(L.mk),
                       0, kNpadSmall, 0);
  // the large class also takes the partitions the four-wave launch deferred;
  // min(K, kLargeGroups) workgroups per slot step through its list
  hipLaunchKernelGGL((kb_partition_large<MODE_PRIMAL, 1, 1>), dim3(L.mlg.n), dim3(kWave), L.lds_big, s, tab,
                     at(L.mlg), 1, kMaxNpad);
  if (e) HIP_TRY(hipEventRecord(e[1], s));
  if (pass) HIP_TRY(hipStreamWaitEvent(s, L.ev_join, 0));
  hipLaunchKernelGGL(kb_finish, dim3(L.m1.n), dim3(1), 0, s, tab, at(L.m1));
  HIP_TRY(hipGetLastError());
  return LG_OK;
}

// One batched iteration of the lane captured into a graph (tables built,
// nothing runs) and instantiated.
int srv_capture(LgServer* S, SrvLane& L) {
  HIP_TRY(hipStreamBeginCapture(S->cap, hipStreamCaptureModeThreadLocal));
  // every captured iteration carries the dual / catch-up pass
  const int st = srv_launch_iteration(S, L, 0, S->cap, true);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(S->cap, &g);
  if (st || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st ? st : fail(LG_ERR_DEVICE, std::string("request server: capture: ") + hipGetErrorString(e));
  }
  if (L.gexec) (void)hipGraphExecDestroy(L.gexec);
  L.gexec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&L.gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    L.gexec = nullptr;
    return fail(LG_ERR_DEVICE, std::string("request server: hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  L.graph_stale = false;
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
  (void)S;
  int st = solve_end(h, h->own_stream);
  if (!st) st = lg_postprocess(h, h->own_stream);
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

void srv_retire(LgServer* S, SrvLane& L, int sl) {
  SrvSlot& slot = S->slots[sl];
  LgHandle* h = slot.h;
  SrvRequest* r = slot.req;
  h->host_ctl = S->h_ctl[sl];
  int st = srv_finish_handle(S, r, h);
  const std::string err = st ? g_err : std::string();
  lg_destroy(h);
  slot = SrvSlot{};
  L.active.erase(std::find(L.active.begin(), L.active.end(), sl));
  L.dirty = true;
  S->resident--;
  srv_complete(S, r, st, err);
}

// Creates the request's handle in slot `sl` of lane `L` (between batches)
// and starts its solve on the lane's stream; returns LG_ERR_MEMORY (request
// untouched) when the pool cannot hold it now.
int srv_admit(LgServer* S, SrvRequest* r, SrvLane& L, int sl) {
  LgHandle* h = nullptr;
  int st = create_from_prep(r->prep, &S->pool, &S->d_ctl[sl], L.stream, &h);
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
    st = lg_solve(h, L.stream);
    if (!st) st = lg_postprocess(h, L.stream);
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
  if ((st = solve_begin(h, L.stream))) {
    const std::string err = g_err;
    lg_destroy(h);
    srv_complete(S, r, st, err);
    return LG_OK;
  }
  S->slots[sl] = SrvSlot{r, h};
  S->h_ctl[sl].iter = 0;
  L.active.push_back(sl);
  L.dirty = true;
  S->resident++;
  return LG_OK;
}

void srv_fail_all(LgServer* S, const std::string& why) {
  for (SrvLane& L : S->lanes) {
    (void)hipStreamSynchronize(L.stream);
    (void)hipStreamSynchronize(L.side);
    L.in_flight = false;
    for (int sl : std::vector<int>(L.active)) {
      SrvSlot& slot = S->slots[sl];
      lg_destroy(slot.h);
      srv_complete(S, slot.req, LG_ERR_STATE, why);
      slot = SrvSlot{};
    }
    L.active.clear();
    L.dirty = true;
  }
  S->resident = 0;
  std::deque<SrvRequest*> q;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    q.swap(S->queue);
  }
  for (SrvRequest* r : q) srv_complete(S, r, LG_ERR_STATE, why);
}

// Requests to admit now, taken out of the queue (under mu).  The queue stays
// in arrival order, so `passed` never grows from its front to its back: while
// the front has been passed over fewer than max_overtakes times every queued
// request has, and a pick adds one to each request ahead of it.  Once the
// front reaches the bound, admission is in arrival order until it is in.
void srv_select(LgServer* S, std::vector<SrvRequest*>& take) {
  std::deque<SrvRequest*>& q = S->queue;
  int free_slots = S->cfg.max_slots - S->resident;
  size_t avail = S->pool.available();
  while (free_slots > 0 && !q.empty()) {
    size_t pick = q.size();
    SrvRequest* f = q.front();
    if (S->cfg.max_overtakes > 0 && f->passed < S->cfg.max_overtakes) {
      for (size_t i = 0; i < q.size(); i++) {
        const SrvRequest* r = q[i];
        if (S->lanes[r->lane].in_flight || r->bytes > avail) continue;
        if (pick == q.size() || r->num_cells < q[pick]->num_cells) pick = i;
      }
    } else if (!S->lanes[f->lane].in_flight && f->bytes <= avail) {
      pick = 0;
    }
    if (pick == q.size()) break;
    for (size_t i = 0; i < pick; i++) q[i]->passed++;
    take.push_back(q[pick]);
    avail -= q[pick]->bytes;
    q.erase(q.begin() + (std::ptrdiff_t)pick);
    free_slots--;
  }
}

// Back into the queue at its arrival position (under mu).
void srv_requeue(LgServer* S, SrvRequest* r) {
  auto pos = std::upper_bound(S->queue.begin(), S->queue.end(), r->ticket,
                              [](int64_t t, const SrvRequest* x) { return t < x->ticket; });
  S->queue.insert(pos, r);
}

// Enqueues one batch of the lane: tables if they changed, iters_per_sync
// batched iterations, the Ctl copy and the lane's event.
int srv_lane_launch(LgServer* S, SrvLane& L) {
  int st = L.dirty ? srv_build_tables(S, L) : LG_OK;
  if (st) return st;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    L.prof_batch = S->prof;
  }
  if (L.prof_batch) {
    // iteration counters at the start of the batch
    L.iter_before.assign(L.top_slot - L.slot_lo, 0);
    for (int sl : L.active) L.iter_before[sl - L.slot_lo] = S->h_ctl[sl].iter;
  }
  const int iters = S->cfg.iters_per_sync;
  if (S->cfg.use_graphs && !L.prof_batch) {
    if (!L.gexec || L.graph_stale) st = srv_capture(S, L);
    for (int it = 0; !st && it < iters; it++)
      if (hipGraphLaunch(L.gexec, L.stream) != hipSuccess) st = fail(LG_ERR_DEVICE, "request server: graph launch");
  } else {
    for (int it = 0; !st && it < iters; it++) {
      // lazy duals: the catch-up pass every kCatchupPeriod batched
      // iterations of the lane (every iteration while one of its slots runs
      // the adaptive search, whose prices must stay current)
      const int64_t n = L.iterations + it;
      const bool pass = L.any_adaptive || n % kCatchupPeriod == kCatchupPeriod - 1;
      st = srv_launch_iteration(S, L, it, L.stream, pass);
    }
  }
  if (!st && hipMemcpyAsync(S->h_ctl + L.slot_lo, S->d_ctl + L.slot_lo, (L.top_slot - L.slot_lo) * sizeof(Ctl),
                            hipMemcpyDeviceToHost, L.stream) != hipSuccess)
    st = fail(LG_ERR_DEVICE, "request server: control copy failed");
  if (!st && hipEventRecord(L.ev_done, L.stream) != hipSuccess)
    st = fail(LG_ERR_DEVICE, "request server: event record failed");
  if (!st) L.in_flight = true;
  return st;
}

// The lane's batch has ended (its event completed): counters, timing, and
// retirement of its finished slots.
int srv_lane_finish(LgServer* S, SrvLane& L) {
  if (hipStreamSynchronize(L.stream) != hipSuccess)
    return fail(LG_ERR_DEVICE, "request server: device error in the batched loop");
  L.in_flight = false;
  const int iters = S->cfg.iters_per_sync;
  L.iterations += iters;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->ctr.batches++;
    S->ctr.iterations += iters;
    if (L.prof_batch) {
      // intervals from the batch's first event (fine) placed on the common
      // time line by that event's offset from prof_ref (coarse)
      float base = 0;
      const bool on_line = hipEventElapsedTime(&base, S->prof_ref, L.ev[0]) == hipSuccess;
      for (int it = 0; it < iters; it++) {
        float ms = 0, t0 = 0, t1 = 0;
        if (hipEventElapsedTime(&ms, L.ev[2 * it], L.ev[2 * it + 1]) == hipSuccess) S->prof_ms += ms;
        if (on_line && hipEventElapsedTime(&t0, L.ev[0], L.ev[2 * it]) == hipSuccess &&
            hipEventElapsedTime(&t1, L.ev[0], L.ev[2 * it + 1]) == hipSuccess)
          S->prof_iv.emplace_back((double)base + t0, (double)base + t1);
      }
      S->prof_launches += iters;
      for (int sl : L.active) {
        const Args& a = S->slots[sl].h->a;
        const double ran = S->h_ctl[sl].iter - L.iter_before[sl - L.slot_lo];
        S->prof_bytes += ran * (12.0 * a.x_size * a.y_size + 24.0 * a.N);
      }
    }
  }
  for (int sl : std::vector<int>(L.active))
    if (S->h_ctl[sl].done || S->h_ctl[sl].error) srv_retire(S, L, sl);
  return LG_OK;
}

bool srv_any_in_flight(const LgServer* S) {
  for (const SrvLane& L : S->lanes)
    if (L.in_flight) return true;
  return false;
}

bool srv_lane_ready(const LgServer* S) {
  for (const SrvLane& L : S->lanes)
    if (!L.in_flight && !L.active.empty()) return true;
  return false;
}

void srv_loop(LgServer* S) {
  (void)hipSetDevice(S->device);
  int64_t seen = 0;  // submissions counted at the last admission
  for (;;) {
    std::vector<SrvRequest*> take;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (!srv_any_in_flight(S))
        S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || srv_lane_ready(S); });
      if (S->stop) break;
      seen = S->ctr.submitted;
      srv_select(S, take);
    }
    for (size_t i = 0; i < take.size(); i++) {
      SrvRequest* r = take[i];
      SrvLane& L = S->lanes[r->lane];
      int sl = L.slot_lo;
      while (S->slots[sl].h) sl++;  // fewer than max_slots resident: a slot is free
      int st = srv_admit(S, r, L, sl);
      if (st == LG_ERR_MEMORY && S->resident == 0 && S->pool.live.empty()) {
        // nothing resident: chunks too small for this request are given
        // back and the pool starts again
        for (SrvLane& M : S->lanes) (void)hipStreamSynchronize(M.stream);
        S->pool.destroy();
        st = srv_admit(S, r, L, sl);
      }
      if (st == LG_ERR_MEMORY && S->resident > 0) {
        // fragmentation: wait for a retirement, keep arrival order
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t j = i; j < take.size(); j++) srv_requeue(S, take[j]);
        break;
      }
      if (st) srv_complete(S, r, st, g_err);
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->ctr.peak_slots = std::max<int32_t>(S->ctr.peak_slots, (int32_t)S->resident);
    }
    int st = LG_OK;
    for (SrvLane& L : S->lanes)
      if (!st && !L.in_flight && !L.active.empty()) st = srv_lane_launch(S, L);
    // wait for the end of a lane's batch, or for a new request while a lane
    // is between batches
    std::vector<SrvLane*> ended;
    while (!st) {
      bool any = false;
      for (SrvLane& L : S->lanes) {
        if (!L.in_flight) continue;
        any = true;
        const hipError_t q = hipEventQuery(L.ev_done);
        if (q == hipSuccess) ended.push_back(&L);
        else if (q != hipErrorNotReady) st = fail(LG_ERR_DEVICE, "request server: device error in the batched loop");
      }
      if (st || !ended.empty() || !any) break;
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->stop) break;
      if (S->ctr.submitted != seen && S->lanes.size() > 1) break;
      S->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
    for (SrvLane* L : ended)
      if (!st) st = srv_lane_finish(S, *L);
    if (st) {
      S->fatal = g_err;
      srv_fail_all(S, S->fatal);
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
      break;
    }
  }
  srv_fail_all(S, S->fatal.empty() ? "request server stopped" : S->fatal);
}

void srv_free(LgServer* S) {
  for (SrvLane& L : S->lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (L.side) (void)hipStreamSynchronize(L.side);
    for (auto& e : L.ev) (void)hipEventDestroy(e);
    if (L.gexec) (void)hipGraphExecDestroy(L.gexec);
    if (L.d_map) (void)hipFree(L.d_map);
    if (L.h_map) (void)hipHostFree(L.h_map);
    if (L.ev_fork) (void)hipEventDestroy(L.ev_fork);
    if (L.ev_join) (void)hipEventDestroy(L.ev_join);
    if (L.ev_done) (void)hipEventDestroy(L.ev_done);
    if (L.side) (void)hipStreamDestroy(L.side);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  if (S->d_ctl) (void)hipFree(S->d_ctl);
  if (S->d_tab) (void)hipFree(S->d_tab);
  if (S->h_ctl) (void)hipHostFree(S->h_ctl);
  if (S->h_tab) (void)hipHostFree(S->h_tab);
  if (S->cap) (void)hipStreamDestroy(S->cap);
  if (S->prof_ref) (void)hipEventDestroy(S->prof_ref);
  if (S->prof_stream) (void)hipStreamDestroy(S->prof_stream);
  delete S;
}

}  // namespace

extern "C" {

void lg_srv_default_config(LgSrvConfig* c) {
  if (!c) return;
  c->max_slots = 128;
  c->iters_per_sync = 16;
  c->mem_budget_bytes = 200ll * 1000 * 1000 * 1000;
  c->pool_chunk_bytes = 256ll << 20;
  c->use_graphs = 0;
  c->short_cells = 20000;
  c->max_overtakes = 64;
  c->reserved = 0;
}

int lg_srv_create(const LgSrvConfig* c, LgServer** out) {
  if (!c || !out) return fail(LG_ERR_ARG, "null argument");
  if (c->max_slots < 1 || c->max_slots > 1024 || c->iters_per_sync < 1 || c->iters_per_sync > 1024 ||
      c->mem_budget_bytes <= 0 || c->pool_chunk_bytes <= 0 || c->short_cells < 0 || c->max_overtakes < 0)
    return fail(LG_ERR_ARG, "bad server configuration");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LG_ERR_DEVICE, "no HIP device");
  int st = upload_directions();
  if (st) return st;
  LgServer* S = new LgServer();
  S->cfg = *c;
  auto bail = [&](int code) {
    std::string msg = g_err;
    srv_free(S);
    g_err = msg;
    return code;
  };
  if (hipGetDevice(&S->device) != hipSuccess) return bail(fail(LG_ERR_DEVICE, "hipGetDevice"));
  S->pool.budget = (size_t)c->mem_budget_bytes;
  S->pool.chunk_bytes = (size_t)c->pool_chunk_bytes;
  const int nl = c->short_cells > 0 ? 2 : 1;
  const size_t ns = (size_t)c->max_slots * nl;
  S->lanes.resize(nl);
  for (int l = 0; l < nl; l++) S->lanes[l].slot_lo = l * c->max_slots;
  // the lanes' main streams first, so that they take distinct hardware queues
  bool ok = true;
  for (SrvLane& L : S->lanes) ok = ok && hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess;
  for (SrvLane& L : S->lanes)
    ok = ok && hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&L.ev_fork, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&L.ev_join, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&L.ev_done, hipEventDisableTiming) == hipSuccess;
  if (!ok || hipStreamCreateWithFlags(&S->cap, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&S->d_ctl, ns * sizeof(Ctl)) != hipSuccess ||
      hipMalloc((void**)&S->d_tab, ns * sizeof(Args)) != hipSuccess ||
      hipHostMalloc((void**)&S->h_ctl, ns * sizeof(Ctl), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&S->h_tab, ns * sizeof(Args), hipHostMallocDefault) != hipSuccess)
    return bail(fail(LG_ERR_MEMORY, "request server: creation of the streams or slot tables failed"));
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
  r->lane = S->lanes.size() > 1 && r->num_cells < S->cfg.short_cells ? 1 : 0;
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
  out->active = (int32_t)S->resident;
  out->queued = (int32_t)S->queue.size();
  out->pool_reserved_bytes = (int64_t)S->pool.reserved;
  out->pool_peak_bytes = (int64_t)S->pool.peak;
  return LG_OK;
}

int lg_srv_profile(LgServer* S, int enable) {
  if (!S) return fail(LG_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  if (enable) {
    for (SrvLane& L : S->lanes)
      if (L.ev.empty()) {
        L.ev.resize(2 * (size_t)S->cfg.iters_per_sync);
        for (auto& e : L.ev) HIP_TRY(hipEventCreate(&e));
      }
    if (!S->prof_stream) HIP_TRY(hipStreamCreateWithFlags(&S->prof_stream, hipStreamNonBlocking));
    if (!S->prof_ref) HIP_TRY(hipEventCreate(&S->prof_ref));
    HIP_TRY(hipEventRecord(S->prof_ref, S->prof_stream));
  }
  S->prof = enable ? 1 : 0;
  S->prof_ms = S->prof_bytes = 0;
  S->prof_launches = 0;
  S->prof_iv.clear();
  return LG_OK;
}

int lg_srv_profile_read(LgServer* S, double* ms, int64_t* iterations, double* bytes, double* ms_union) {
  if (!S) return fail(LG_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  if (ms) *ms = S->prof_ms;
  if (iterations) *iterations = S->prof_launches;
  if (bytes) *bytes = S->prof_bytes;
  if (ms_union) {
    std::vector<std::pair<double, double>> iv = S->prof_iv;
    std::sort(iv.begin(), iv.end());
    double u = 0, lo = 0, hi = 0;
    bool open = false;
    for (const auto& x : iv) {
      if (open && x.first <= hi) {
        hi = std::max(hi, x.second);
        continue;
      }
      if (open) u += hi - lo;
      lo = x.first;
      hi = x.second;
      open = true;
    }
    if (open) u += hi - lo;
    *ms_union = u;
  }
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
  srv_free(S);
  return LG_OK;
}

}  // extern "C"
