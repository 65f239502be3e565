// Device layer of the legalizer: C ABI lg_create / lg_solve / lg_postprocess /
// lg_check / lg_copy_out / lg_destroy (include/skyline_lg.h).  Replaces the
// reference's CudaDatabase + CudaAdmmLegalizer + CudaPostProcess
// (database/CudaDatabase.h:16-49, admm_legalizer/CudaAdmmLegalizer.h:19-33,
// post_process/CudaPostProcess.h:12-20).  The handle owns all device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <cstdlib>
#include <random>
#include <chrono>
#include <string>
#include <thread>
#include <vector>
#include <map>
#include <unordered_map>

#include "../../include/skyline_lg.h"
#include "lg_common.h"
#include "lg_internal.h"
#include "lg_grid.h"

// The kernels and this orchestration form one translation unit, so the
// templated partition kernel and the __constant__ direction tables are local.
#include "lg_kernels.hip"

using namespace lgdev;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(LG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(LG_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  return LG_OK;
}

bool g_dirs_uploaded[64] = {false};

int upload_directions() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 64 && g_dirs_uploaded[dev]) return LG_OK;
  std::vector<int2> dd, dw;
  for (int y = -6; y <= 6; y++) {  // Legalizer::makeDirectionVectors (Legalizer.cpp:252-275)
    static const int xr[7] = {40, 20, 10, 5, 2, 1, 0};
    for (int x = -xr[std::abs(y)]; x <= xr[std::abs(y)]; x++) dd.push_back(make_int2(x, y));
  }
  for (int y = -12; y <= 12; y++) {  // (Legalizer.cpp:323-351)
    static const int xr[13] = {280, 240, 200, 160, 120, 80, 40, 20, 10, 4, 2, 1, 0};
    for (int x = -xr[std::abs(y)]; x <= xr[std::abs(y)]; x++) dw.push_back(make_int2(x, y));
  }
  if ((int)dd.size() != kNumDirDefault || (int)dw.size() != kNumDirWide)
    return fail(LG_ERR_STATE, "direction table size");
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_dir_default), dd.data(), dd.size() * sizeof(int2)));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_dir_wide), dw.data(), dw.size() * sizeof(int2)));
  if (dev < 64) g_dirs_uploaded[dev] = true;
  return LG_OK;
}

}  // namespace

// Device-memory pool of the request server: chunks of device memory taken
// with hipMalloc on demand up to a byte budget, first-fit sub-allocation with
// coalescing of free neighbours.  Requests are admitted only while their
// buffers fit the budget; nothing is returned to the runtime before the
// server ends (hipFree would synchronise the device mid-stream).
struct DevPool {
  struct Chunk {
    char* base;
    size_t cap;
    std::map<size_t, size_t> free;  // offset -> bytes
  };
  size_t budget = 0, chunk_bytes = 256u << 20, reserved = 0, in_use = 0, peak = 0;
  std::vector<Chunk> chunks;
  std::unordered_map<void*, std::pair<int, size_t>> live;  // block -> (chunk, bytes)

  static size_t round(size_t n) { return (std::max<size_t>(n, 1) + 255) & ~(size_t)255; }

  void* take(int ci, size_t n) {
    Chunk& c = chunks[ci];
    for (auto it = c.free.begin(); it != c.free.end(); ++it) {
      if (it->second < n) continue;
      const size_t off = it->first, rest = it->second - n;
      c.free.erase(it);
      if (rest) c.free.emplace(off + n, rest);
      void* p = c.base + off;
      live.emplace(p, std::make_pair(ci, n));
      in_use += n;
      peak = std::max(peak, in_use);
      return p;
    }
    return nullptr;
  }
  // nullptr when the budget is exhausted (or the runtime refuses a chunk)
  void* alloc(size_t n) {
    n = round(n);
    for (int ci = 0; ci < (int)chunks.size(); ci++)
      if (void* p = take(ci, n)) return p;
    if (reserved + n > budget) return nullptr;
    const size_t want = std::max(std::min(chunk_bytes, budget - reserved), n);
    char* base = nullptr;
    if (hipMalloc((void**)&base, want) != hipSuccess) return nullptr;
    reserved += want;
    chunks.push_back(Chunk{base, want, {}});
    chunks.back().free.emplace(0, want);
    return take((int)chunks.size() - 1, n);
  }
  void release(void* p) {
    auto it = live.find(p);
    if (it == live.end()) return;
    Chunk& c = chunks[it->second.first];
    size_t off = (char*)p - c.base, n = it->second.second;
    in_use -= n;
    live.erase(it);
    auto nx = c.free.lower_bound(off);
    if (nx != c.free.end() && off + n == nx->first) {  // merge the right neighbour
      n += nx->second;
      nx = c.free.erase(nx);
    }
    if (nx != c.free.begin()) {  // merge the left neighbour
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += n;
        return;
      }
    }
    c.free.emplace(off, n);
  }
  size_t available() const {
    size_t f = budget - reserved;
    for (const Chunk& c : chunks)
      for (const auto& kv : c.free) f += kv.second;
    return f;
  }
  void destroy() {
    for (Chunk& c : chunks) (void)hipFree(c.base);
    chunks.clear();
    live.clear();
    reserved = in_use = 0;
  }
};

// Host-side preparation of one legalization (what lg_create computes before
// any device call): validated sizes, the pixel codes and region ids in the
// device's row-major order, grid positions, packed cell codes, the height
// histogram, random sorting priorities and the fold-offset maps.
struct HostPrep {
  LgParams p;
  int x_size = 0, y_size = 0, site_width = 0, row_height = 0, core_lx = 0, core_ly = 0;
  int ncol = 0, nrow = 0, K = 0, N = 0;
  size_t P = 0;
  std::vector<int16_t> pcode;
  void* dev_pcode = nullptr;  // pixel codes built on the device (lg_create_device_grid)
  std::vector<int32_t> region_rm, vdd, rx0, ry0, rx1, ry1;
  std::vector<int32_t> lxg, lyg, w, h, code, gnd, ox, oy, rprio, khn, colmap, rowmap;
  // dual segments: maximal runs of columns (rows) inside one partition
  // column (row) under all three fold offsets; segment -> first x (y)
  std::vector<int32_t> xseg, yseg, xseg0, yseg0;
  // ICCAD 2017 technology (tech): power metal per pixel (row-major), edge
  // types per cell (left | right << 8), master per cell, pins {bgn, end,
  // layer, 0} per master, the rule table (indices 0..4)
  bool tech = false;
  std::vector<int8_t> pmetal;
  std::vector<int32_t> cedge, cmacro, mpin, pins;
  int32_t rules[5] = {0, 0, 0, 0, 0};
};

// Segments of [0, n) cut wherever the partition column (row) changes under
// any fold offset (x + hint / 3 * tf a multiple of hint): a segment lies in
// one partition in every iteration.
// The ids count three segments per period of `hint` (an empty one where two
// cuts coincide), as the kernels' seg_axis computes them.
void dual_segments(int n, int hint, std::vector<int32_t>& seg, std::vector<int32_t>& first) {
  seg.resize(n);
  const int t = hint / 3;
  for (int x = 0; x < n; x++) {
    const int q = x / hint, r = x % hint;
    seg[x] = 3 * q + (r >= hint - 2 * t) + (r >= hint - t);
  }
  const int ns = n > 0 ? seg[n - 1] + 1 : 0;
  first.assign(ns, -1);
  for (int x = n - 1; x >= 0; x--) first[seg[x]] = x;
  // an empty id gets the next segment's first column (its stamp is unused)
  int next = n > 0 ? n - 1 : 0;
  for (int s = ns - 1; s >= 0; s--) {
    if (first[s] < 0) first[s] = next;
    next = first[s];
  }
}

struct LgHandle {
  LgParams p;
  Args a;
  hipStream_t own_stream = nullptr;
  bool owns_stream = true;     // false: the stream belongs to a request server
  DevPool* pool = nullptr;     // non-null: buffers come from a server's pool,
  char* blk = nullptr;         //   carved from one block of the pool
  size_t blk_size = 0, blk_used = 0;
  int solved = 0, posted = 0;
  size_t P = 0;
  // device buffers (freed in lg_destroy)
  std::vector<void*> bufs;
  Ctl host_ctl;
  std::vector<int32_t> trace;
  int trace_n = 0;
  LgVioStat vio{};
  int checked = 0;
  float coeff_init_rho, coeff_init_dual;
  int32_t *init_lx = nullptr, *init_ly = nullptr;  // positions given at lg_create
  // partition-kernel timing (lg_profile_partition): three events per
  // iteration bracket the two LDS-class launches
  int prof = 0;
  std::vector<hipEvent_t> ev;
  int ev_used = 0;
  double prof_ms[2] = {0, 0};
  int64_t prof_n[2] = {0, 0};
  // side stream of the dual pass, which runs beside the partition kernels
  // (it updates only the pixels of partitions they leave alone); created at
  // the first solve
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // HIP graph of kGraphIters iterations (lg_use_graphs), captured on `cap`
  // at the first solve and replayed on the solve's stream
  int graphs = 0;
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
};

constexpr int kGraphIters = 8;
constexpr size_t kSideMinPixels = 8u << 20;
// lazy duals: a catch-up pass every kCatchupPeriod iterations bounds how many
// dual updates a staged price may have to catch up on
constexpr int kCatchupPeriod = 32;

namespace {

// device buffer of a handle: carved from its pool block (else from the
// pool itself), or hipMalloc for a handle outside a server
template <typename T>
int halloc(LgHandle* h, T** p, size_t n) {
  if (h->pool) {
    const size_t b = DevPool::round(std::max<size_t>(n, 1) * sizeof(T));
    if (h->blk && h->blk_used + b <= h->blk_size) {
      *p = (T*)(h->blk + h->blk_used);
      h->blk_used += b;
      return LG_OK;
    }
    *p = (T*)h->pool->alloc(b);
    if (!*p) return fail(LG_ERR_MEMORY, "request server: device-memory budget exhausted");
    h->bufs.push_back((void*)*p);
    return LG_OK;
  }
  int st = dalloc(p, n);
  if (st == LG_OK) h->bufs.push_back((void*)*p);
  return st;
}

// temporary device buffer of a handle (returned with hfree_tmp)
template <typename T>
int halloc_tmp(LgHandle* h, T** p, size_t n) {
  if (h->pool) {
    *p = (T*)h->pool->alloc(std::max<size_t>(n, 1) * sizeof(T));
    return *p ? LG_OK : fail(LG_ERR_MEMORY, "request server: device-memory budget exhausted");
  }
  return dalloc(p, n);
}

void hfree_tmp(LgHandle* h, void* p) {
  if (h->pool) h->pool->release(p);
  else (void)hipFree(p);
}

template <typename T>
int upload(LgHandle* h, T** dst, const T* src, size_t n) {
  int st = halloc(h, dst, n);
  if (st) return st;
  if (n) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return LG_OK;
}

hipStream_t pick(LgHandle* h, void* s) { return s ? (hipStream_t)s : h->own_stream; }

size_t sort_bytes_per(int algo) {
  return (algo == LG_SORT_S_MERGE || algo == LG_SORT_M_MERGE) ? 16 : (algo == LG_SORT_CUB ? 12 : 8);
}

// LDS of one partition workgroup: pixel tile {usage|code, dual} + dual/rho +
// vdd rows, plus the sort buffers when the sorts run in LDS.
size_t partition_lds_bytes(int x_hint, int y_hint, int algo, int npad_hi, int lds_sort, bool tech) {
  // three tile arrays (usage|code, penalty term, dual) with the kernel's row
  // stride, plus the vdd rows
  const size_t rs = x_hint <= kRowStride ? kRowStride : x_hint;
  const size_t tsz = (rs * y_hint + 3) & ~(size_t)3;
  size_t tile = 3 * tsz * 4 + (size_t)((y_hint + 3) & ~3) * 4;
  // with technology: the left / right edge maps (4 bytes each) and the
  // power metal (1 byte) of every tile pixel
  if (tech) tile += tsz * 9;
  // + the four-wave launches' exchange words (64) and sort arrays (2 x 256)
  return tile + (lds_sort ? (size_t)npad_hi * sort_bytes_per(algo) : (64 + 512) * 4);
}

size_t partition_lds_bytes(const LgHandle* h, int npad_hi, int lds_sort) {
  return partition_lds_bytes(h->a.x_hint, h->a.y_hint, h->p.sorting_algorithm, npad_hi, lds_sort, h->a.tech != 0);
}

// LDS of a workgroup on gfx950
constexpr size_t kMaxLds = 160 * 1024;

// Launch one partition-kernel pass.  Two LDS size classes; the device picks
// the class from ctl->npad and the other launch exits at once.  Up to 1024
// slots the default bitonic order is sorted in registers.
template <int MODE>
int launch_partition(LgHandle* h, hipStream_t s) {
  // the pin-access and edge-spacing terms are evaluated by the one-wave
  // launches only (their LDS tile holds the edge maps and power metal)
  const int reg_ok = !h->a.tech &&
                     (h->p.sorting_algorithm == LG_SORT_M_BITONIC || h->p.sorting_algorithm == LG_SORT_S_BITONIC);
  const bool prof = h->prof && MODE == MODE_PRIMAL && h->ev_used + 3 <= (int)h->ev.size();
  hipEvent_t* e = prof ? &h->ev[h->ev_used] : nullptr;
  if (prof) HIP_TRY(hipEventRecord(e[0], s));
  // small class: register sorts when the algorithm is a bitonic network
  size_t lds0 = partition_lds_bytes(h, kNpadSmall, reg_ok ? 0 : 1);
  if (reg_ok) {
    if constexpr (MODE == MODE_PRIMAL)
      hipLaunchKernelGGL((k_partition<MODE_PRIMAL, 0, kPartWaves>), dim3(h->a.K), dim3(kPartWaves * kWave), lds0, s,
                         h->a, 0, kNpadSmall);
    else
      hipLaunchKernelGGL((k_partition<MODE, 0, 1>), dim3(h->a.K), dim3(kWave), lds0, s, h->a, 0, kNpadSmall);
  } else
    hipLaunchKernelGGL((k_partition<MODE, 1, 1>), dim3(h->a.K), dim3(kWave), lds0, s, h->a, 0, kNpadSmall);
  if (prof) HIP_TRY(hipEventRecord(e[1], s));
  // sort buffers of up to kMaxNpadLds slots in LDS; larger partitions (up
  // to kMaxNpad) sort in Args::huge
  size_t lds1 = partition_lds_bytes(h, kMaxNpadLds, 1);
  // the primal step's large class also takes the partitions the four-wave
  // launch deferred (its list, every size)
  if constexpr (MODE == MODE_PRIMAL) {
    // min(K, 1024) workgroups step through the list, which is short
    hipLaunchKernelGGL((k_partition_large<MODE, 1, 1>), dim3(std::min(h->a.K, 1024)), dim3(kWave), lds1, s, h->a, 1,
                       kMaxNpad);
  } else {
    hipLaunchKernelGGL((k_partition<MODE, 1, 1>), dim3(h->a.K), dim3(kWave), lds1, s, h->a, kNpadSmall + 1,
                       kMaxNpad);
  }
  if (prof) {
    HIP_TRY(hipEventRecord(e[2], s));
    h->ev_used += 3;
  }
  HIP_TRY(hipGetLastError());
  return LG_OK;
}

// Fold the recorded partition-launch events into the totals (stream synced).
int collect_profile(LgHandle* h) {
  for (int i = 0; i + 3 <= h->ev_used; i += 3) {
    for (int c = 0; c < 2; c++) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, h->ev[i + c], h->ev[i + c + 1]));
      h->prof_ms[c] += ms;
      h->prof_n[c]++;
    }
  }
  h->ev_used = 0;
  return LG_OK;
}

int read_ctl(LgHandle* h, hipStream_t s) {
  HIP_TRY(hipMemcpyAsync(&h->host_ctl, h->a.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return h->prof ? collect_profile(h) : LG_OK;
}


// The optional ICCAD 2017 technology of lg_create: validated and laid out
// for the device (power metal row-major as int8, edge types packed per
// cell).  Absent parts are zero: no power metal, edge type 0, no pins.
int prepare_tech(const LgGridDesc* g, const LgCellsSoA* c, HostPrep& hp) {
  hp.tech = g->pixel_power_layer || g->num_edge_rules > 0 || c->ledge_type || c->redge_type || c->macro_id;
  if (!hp.tech) return LG_OK;
  const int XS = g->x_size, YS = g->y_size, N = c->num_cells;
  if (g->num_edge_rules < 0 || (g->num_edge_rules > 0 && !g->edge_spacing_rules))
    return fail(LG_ERR_ARG, "edge_spacing_rules missing");
  // types are 0..2, so the kernels read indices up to 4; entries the table
  // does not hold are 0
  for (int i = 0; i < std::min(g->num_edge_rules, 5); i++) {
    if (g->edge_spacing_rules[i] < 0 || g->edge_spacing_rules[i] > 1024)
      return fail(LG_ERR_ARG, "edge spacing outside [0, 1024] sites");
    hp.rules[i] = g->edge_spacing_rules[i];
  }
  hp.pmetal.assign((size_t)XS * YS, (int8_t)-1);
  if (g->pixel_power_layer)
    for (int x = 0; x < XS; x++)
      for (int y = 0; y < YS; y++) {
        const int l = g->pixel_power_layer[(size_t)x * YS + y];
        if (l < -1 || l > 127) return fail(LG_ERR_ARG, "power-metal layer index outside [-1, 127]");
        hp.pmetal[(size_t)y * XS + x] = (int8_t)l;
      }
  hp.cedge.assign(N, 0);
  for (int i = 0; i < N; i++) {
    const int lt = c->ledge_type ? c->ledge_type[i] : 0, rt = c->redge_type ? c->redge_type[i] : 0;
    if (lt < 0 || lt > 2 || rt < 0 || rt > 2) return fail(LG_ERR_ARG, "edge type outside {0, 1, 2}");
    hp.cedge[i] = lt | rt << 8;
  }
  if (c->macro_id) {
    const int M = c->num_macros;
    if (M <= 0 || !c->macro_pin_start) return fail(LG_ERR_ARG, "macro_id given without masters");
    hp.mpin.assign(c->macro_pin_start, c->macro_pin_start + M + 1);
    if (hp.mpin[0] != 0) return fail(LG_ERR_ARG, "macro_pin_start[0] must be 0");
    for (int m = 0; m < M; m++)
      if (hp.mpin[m + 1] < hp.mpin[m]) return fail(LG_ERR_ARG, "macro_pin_start not non-decreasing");
    const int np = hp.mpin[M];
    if (np > 0 && (!c->pin_bgn || !c->pin_end || !c->pin_layer)) return fail(LG_ERR_ARG, "pin arrays missing");
    hp.pins.resize(4 * (size_t)np);
    for (int i = 0; i < np; i++) {
      hp.pins[4 * i] = c->pin_bgn[i];
      hp.pins[4 * i + 1] = c->pin_end[i];
      hp.pins[4 * i + 2] = c->pin_layer[i];
      hp.pins[4 * i + 3] = 0;
    }
    hp.cmacro.assign(c->macro_id, c->macro_id + N);
    for (int i = 0; i < N; i++) {
      const int m = hp.cmacro[i];
      if (m < 0 || m >= M) return fail(LG_ERR_ARG, "macro_id outside [0, num_macros)");
      const int cells = c->w_grid[i] * c->h_grid[i];
      for (int k = hp.mpin[m]; k < hp.mpin[m + 1]; k++)  // the pin box lies in the cell
        if (c->pin_bgn[k] < 0 || c->pin_bgn[k] >= cells || c->pin_end[k] < 0 || c->pin_end[k] >= cells)
          return fail(LG_ERR_ARG, "pin box outside its cell");
    }
  }
  return LG_OK;
}

// Validation and host-side preparation of lg_create (no device calls).
int prepare_host(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, HostPrep& hp,
                 bool pixels_on_device = false) {
  if (!g || !c || !p) return fail(LG_ERR_ARG, "null argument");
  if (g->x_size <= 0 || g->y_size <= 0 || c->num_cells < 0) return fail(LG_ERR_ARG, "bad sizes");
  if (g->site_width <= 0 || g->row_height <= 0) return fail(LG_ERR_ARG, "bad site/row size");
  if (p->x_hint < 3 || p->y_hint < 3 || p->iter_update_partition <= 0 || p->iter_update_dual_step <= 0)
    return fail(LG_ERR_ARG, "bad partition / schedule parameters");
  if (g->x_size > 65535 || g->y_size > 65535) return fail(LG_ERR_ARG, "grid larger than 65535 sites/rows");
  if ((long long)g->x_size * g->y_size >= (1ll << 30)) return fail(LG_ERR_ARG, "grid of 2^30 pixels or more");
  if (p->use_adaptive_search && (p->stagnation_window < 1 || p->stagnation_window > kMaxStagnationWindow))
    return fail(LG_ERR_ARG, "stagnation_window outside [1, 64]");
  hp.p = *p;
  hp.x_size = g->x_size; hp.y_size = g->y_size;
  hp.site_width = g->site_width; hp.row_height = g->row_height;
  hp.core_lx = g->core_lx; hp.core_ly = g->core_ly;
  // CudaAdmmLegalizer ctor (CudaAdmmLegalizer.cu:86-89)
  hp.ncol = g->x_size % p->x_hint == 0 ? g->x_size / p->x_hint : g->x_size / p->x_hint + 1;
  hp.nrow = g->y_size % p->y_hint == 0 ? g->y_size / p->y_hint : g->y_size / p->y_hint + 1;
  hp.K = hp.ncol * hp.nrow;
  hp.N = c->num_cells;
  const size_t P = (size_t)g->x_size * g->y_size;
  hp.P = P;
  const int N = hp.N;
  // pixel code: valid ? group + 2 : 0.  The ABI's pixel arrays are x-major
  // (x * y_size + y, the reference's Grid layout); the device keeps them
  // row-major (y * x_size + x).
  if (!pixels_on_device) hp.pcode.assign(P, 0);
  // pixel regions are read by the adaptive search only
  const bool regions = p->use_adaptive_search && g->pixel_region;
  if (p->use_adaptive_search) hp.region_rm.assign(P, -1);
  const int XS = g->x_size, YS = g->y_size;
  // transposed in 64 x 64 blocks (cache-friendly on both sides), column
  // blocks spread over host threads
  const int nxb = (XS + 63) / 64;
  const int nth = std::max(1, std::min({p->num_cpu_threads, nxb, 32}));
  std::vector<int> bad(nth, 0), bad_region(nth, 0);
  const int nreg = std::max(g->num_regions, 0);
  auto work = [&](int t) {
    for (int xb = t; xb < nxb; xb += nth)
      for (int y0 = 0; y0 < YS; y0 += 64)
        for (int x = xb * 64; x < std::min(xb * 64 + 64, XS); x++)
          for (int y = y0; y < std::min(y0 + 64, YS); y++) {
            size_t i = (size_t)x * YS + y, j = (size_t)y * XS + x;
            int grp = g->pixel_group[i];
            if (g->pixel_valid[i] && (grp + 2 > 0x7fff || grp + 2 < 0)) bad[t] = 1;
            hp.pcode[j] = g->pixel_valid[i] ? (int16_t)(grp + 2) : (int16_t)0;
            if (regions) {
              const int rid = g->pixel_region[i];
              if (rid < -1 || rid >= nreg) bad_region[t] = 1;
              hp.region_rm[j] = rid;
            }
          }
  };
  if (pixels_on_device) {
    if (p->use_adaptive_search) return fail(LG_ERR_ARG, "the adaptive search reads host pixel regions");
  } else if (nth == 1 || P < (1u << 20)) {
    for (int t = 0; t < nth; t++) work(t);
  } else {
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
  }
  for (int t = 0; t < nth; t++) {
    if (bad[t]) return fail(LG_ERR_ARG, "fence group index out of range");
    if (bad_region[t]) return fail(LG_ERR_ARG, "pixel region id outside [-1, num_regions)");
  }
  hp.vdd.assign(g->row_vdd_up, g->row_vdd_up + YS);
  const int nr = std::max(g->num_regions, 0);
  hp.rx0.assign(g->region_x_min, g->region_x_min + nr);
  hp.ry0.assign(g->region_y_min, g->region_y_min + nr);
  hp.rx1.assign(g->region_x_max, g->region_x_max + nr);
  hp.ry1.assign(g->region_y_max, g->region_y_max + nr);
  hp.lxg.resize(N); hp.lyg.resize(N); hp.code.resize(N);
  int max_h = 1;
  for (int i = 0; i < N; i++) {
    // dbu_to_grid_functor (DatabaseKernels.hpp:7-20)
    hp.lxg[i] = (c->lx_dbu[i] - g->core_lx) / g->site_width;
    hp.lyg[i] = (c->ly_dbu[i] - g->core_ly) / g->row_height;
    hp.code[i] = c->group[i] + 2;
    int w = c->w_grid[i], hh = c->h_grid[i];
    if (w <= 0 || hh <= 0 || w > 4095 || hh > 63 || hp.code[i] <= 0 || hp.code[i] > 8191)
      return fail(LG_ERR_ARG, "cell size or group outside the supported range");
    // A cell wider (taller) than a partition spans partitions of different
    // colours and, as in the reference (assignPartitionKernel,
    // AdmmKernels.hpp:150-161), belongs to none: the primal step and the
    // final refine never move it.  Only when its corner columns can fall on
    // the same colour through the wrap-around of the colouring (a footprint
    // across all ncol columns) would the reference hand it to a partition
    // whose tile does not hold it; such cells are refused.
    if ((hp.ncol > 1 && w >= (hp.ncol - 1) * p->x_hint + 2) || (hp.nrow > 1 && hh >= (hp.nrow - 1) * p->y_hint + 2))
      return fail(LG_ERR_DESIGN, "cell spanning every partition column (row): wider than (ncol-1)*x_hint+1 sites");
    if (hp.lxg[i] < 0 || hp.lyg[i] < 0 || hp.lxg[i] + w > g->x_size || hp.lyg[i] + hh > g->y_size)
      return fail(LG_ERR_ARG, "cell outside the grid (not preplaced?)");
    max_h = std::max(max_h, hh);
  }
  hp.w.assign(c->w_grid, c->w_grid + N);
  hp.h.assign(c->h_grid, c->h_grid + N);
  hp.gnd.assign(c->ground_at_bottom, c->ground_at_bottom + N);
  hp.ox.assign(c->orig_lx_dbu, c->orig_lx_dbu + N);
  hp.oy.assign(c->orig_ly_dbu, c->orig_ly_dbu + N);
  hp.khn.assign(max_h, 0);
  for (int i = 0; i < N; i++) hp.khn[c->h_grid[i] - 1]++;
  hp.rprio.assign(N, 0);
  if (p->sorting_policy == LG_SORT_POLICY_RANDOM) {  // CudaAdmmLegalizer.cu:120-130
    std::iota(hp.rprio.begin(), hp.rprio.end(), 0);
    std::mt19937 gen(p->random_sorting_seed);
    std::shuffle(hp.rprio.begin(), hp.rprio.end(), gen);
  }
  // partition column / row of every x / y for the three fold offsets
  // (paintPixelsKernel, AdmmKernels.hpp:164-191)
  hp.colmap.resize(3 * (size_t)XS);
  hp.rowmap.resize(3 * (size_t)YS);
  for (int tf = 0; tf < 3; tf++) {
    const int ox = p->x_hint / 3 * tf, oy = p->y_hint / 3 * tf;
    for (int x = 0; x < XS; x++) hp.colmap[tf * (size_t)XS + x] = ((x + ox) / p->x_hint) % hp.ncol;
    for (int y = 0; y < YS; y++) hp.rowmap[tf * (size_t)YS + y] = ((y + oy) / p->y_hint) % hp.nrow;
  }
  dual_segments(XS, p->x_hint, hp.xseg, hp.xseg0);
  dual_segments(YS, p->y_hint, hp.yseg, hp.yseg0);
  return prepare_tech(g, c, hp);
}

// Partitions of one step that can sort above kMaxNpadLds slots at once:
// each holds more than kMaxNpadLds of the N cells.
int huge_slots(int N, int K) { return std::max(1, std::min(K, N / (kMaxNpadLds + 1))); }

// Device bytes a handle of `hp` takes (an upper bound of create_from_prep's
// buffers, each rounded to the pool's 256-byte blocks).
size_t handle_bytes(const HostPrep& hp) {
  const size_t P = hp.P, N = std::max(hp.N, 1), K = hp.K, blk = 256;
  size_t b = P * (2 + (hp.p.use_adaptive_search ? 4 : 0) + 4 + 4);  // pcode, region, usage, dual
  b += N * 4 * (2 + 2 + 8 + 1 + 1 + 2 + 2);           // positions, cells, rot, grouped, flags, timing
  b += N * 4 * 4 * 2;                                 // cellrec, rec_b
  b += K * 4 * ((size_t)kMaxNpad + 1 + 1 + 2 + 1 + 1 + 2 + 2);  // buckets, counts, offsets, flags, lists
  b += (3 * (size_t)hp.x_size + 3 * (size_t)hp.y_size + hp.y_size + 4 * hp.rx0.size()) * 4;
  b += (size_t)std::max(hp.p.max_admm_iter, 1) * 4 + sizeof(Ctl);
  b += (size_t)huge_slots(hp.N, hp.K) * 4 * kMaxNpad * 4;  // sort buffers above kMaxNpadLds slots
  // dual segments: tables, stamps, step history
  b += (2 * (size_t)hp.x_size + 2 * (size_t)hp.y_size + hp.xseg0.size() * hp.yseg0.size()) * 4 + 256 * 4;
  b += (size_t)std::max(hp.p.max_admm_iter, 1) * 4;
  if (hp.tech) b += P * (1 + 4 * 4) + N * 8 + (hp.mpin.size() + hp.pins.size()) * 4;  // power metal, edge maps
  return b + 64 * blk;
}

// Device side of lg_create from a prepared design.  With a pool, every
// buffer comes from it; `ext_ctl` places the iteration control in a server's
// Ctl array; `stream` (if given) is the handle's stream, not owned.
int create_from_prep(const HostPrep& hp, DevPool* pool, Ctl* ext_ctl, hipStream_t stream, LgHandle** out) {
  // until the handle exists, a failure frees the device-built pixel codes
  auto early = [&](int code) {
    if (hp.dev_pcode) (void)hipFree(hp.dev_pcode);
    return code;
  };
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return early(fail(LG_ERR_DEVICE, "no HIP device"));
  int st = upload_directions();
  if (st) return early(st);
  const LgParams* p = &hp.p;
  {
    // a partition's pixels are staged in LDS (the large size class also
    // holds its sort buffers there)
    const size_t need = partition_lds_bytes(p->x_hint, p->y_hint, p->sorting_algorithm, kMaxNpadLds, 1, hp.tech);
    if (need > kMaxLds)
      return early(fail(LG_ERR_ARG, "partition of x_hint x y_hint pixels needs " + std::to_string(need) +
                                  " bytes of LDS, more than the " + std::to_string(kMaxLds) + " of a workgroup"));
  }
  if (pool && !(pool->budget >= handle_bytes(hp)))
    return early(fail(LG_ERR_MEMORY, "request larger than the server's device-memory budget"));
  char* blk = nullptr;
  if (pool && !(blk = (char*)pool->alloc(handle_bytes(hp))))
    return early(fail(LG_ERR_MEMORY, "request server: device-memory budget exhausted"));
  LgHandle* h = new LgHandle();
  if (hp.dev_pcode) h->bufs.push_back(hp.dev_pcode);  // freed by lg_destroy from here on
  h->p = *p;
  h->pool = pool;
  h->blk = blk;
  h->blk_size = blk ? handle_bytes(hp) : 0;
  h->coeff_init_rho = p->coeff_to_init_rho;
  h->coeff_init_dual = p->coeff_to_init_dual;
  Args& a = h->a;
  std::memset(&a, 0, sizeof(a));
  a.x_size = hp.x_size; a.y_size = hp.y_size;
  a.site_width = hp.site_width; a.row_height = hp.row_height;
  a.core_lx = hp.core_lx; a.core_ly = hp.core_ly;
  a.x_hint = p->x_hint; a.y_hint = p->y_hint;
  a.ncol = hp.ncol; a.nrow = hp.nrow; a.K = hp.K;
  a.N = hp.N;
  const size_t P = hp.P;
  h->P = P;
  const int N = a.N;

  auto bail = [&](int code) {
    std::string msg = g_err;
    lg_destroy(h);
    g_err = msg;
    return code;
  };

  if (stream) {
    h->own_stream = stream;
    h->owns_stream = false;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking));
  }
  int16_t* d_pcode;
  int32_t *d_region = nullptr, *d_vdd, *d_rx0, *d_ry0, *d_rx1, *d_ry1;
  const size_t nr = hp.rx0.size();
  if (hp.dev_pcode) {  // built on the device (owned by the handle, above)
    d_pcode = (int16_t*)hp.dev_pcode;
  } else if ((st = upload(h, &d_pcode, hp.pcode.data(), P))) {
    return bail(st);
  }
  if (!hp.region_rm.empty() && (st = upload(h, &d_region, hp.region_rm.data(), P))) return bail(st);
  if ((st = upload(h, &d_vdd, hp.vdd.data(), (size_t)hp.y_size))) return bail(st);
  if ((st = upload(h, &d_rx0, hp.rx0.data(), nr))) return bail(st);
  if ((st = upload(h, &d_ry0, hp.ry0.data(), nr))) return bail(st);
  if ((st = upload(h, &d_rx1, hp.rx1.data(), nr))) return bail(st);
  if ((st = upload(h, &d_ry1, hp.ry1.data(), nr))) return bail(st);
  a.pcode = d_pcode; a.pregion = d_region; a.vdd_up = d_vdd;
  a.rxmin = d_rx0; a.rymin = d_ry0; a.rxmax = d_rx1; a.rymax = d_ry1;
  if ((st = halloc(h, &a.usage, P))) return bail(st);
  if ((st = halloc(h, &a.dual, P))) return bail(st);

  int32_t *d_w, *d_h, *d_code, *d_gnd, *d_ox, *d_oy, *d_rp, *d_khn;
  if ((st = upload(h, &h->init_lx, hp.lxg.data(), N))) return bail(st);
  if ((st = upload(h, &h->init_ly, hp.lyg.data(), N))) return bail(st);
  if ((st = halloc(h, &a.lx, N))) return bail(st);
  if ((st = halloc(h, &a.ly, N))) return bail(st);
  if ((st = upload(h, &d_w, hp.w.data(), N))) return bail(st);
  if ((st = upload(h, &d_h, hp.h.data(), N))) return bail(st);
  if ((st = upload(h, &d_code, hp.code.data(), N))) return bail(st);
  if ((st = upload(h, &d_gnd, hp.gnd.data(), N))) return bail(st);
  if ((st = upload(h, &d_ox, hp.ox.data(), N))) return bail(st);
  if ((st = upload(h, &d_oy, hp.oy.data(), N))) return bail(st);
  if ((st = upload(h, &d_rp, hp.rprio.data(), N))) return bail(st);
  if ((st = upload(h, &d_khn, hp.khn.data(), hp.khn.size()))) return bail(st);
  a.w = d_w; a.h = d_h; a.code = d_code; a.gnd = d_gnd; a.ox = d_ox; a.oy = d_oy;
  a.rand_prio = d_rp; a.k_height_num = d_khn;
  if ((st = halloc(h, &a.need_rot, N))) return bail(st);
  HIP_TRY(hipMemset(a.need_rot, 0, std::max(N, 1) * sizeof(int32_t)));
  if ((st = halloc(h, &a.bucket, (size_t)a.K * kMaxNpad))) return bail(st);
  if ((st = halloc(h, &a.bcnt, a.K))) return bail(st);
  HIP_TRY(hipMemset(a.bcnt, 0, std::max(a.K, 1) * sizeof(int32_t)));
  if ((st = halloc(h, &a.pcnt, a.K))) return bail(st);
  if ((st = halloc(h, &a.off, a.K + 2))) return bail(st);
  if ((st = halloc(h, &a.scratch, a.K + 1))) return bail(st);
  if ((st = halloc(h, &a.grouped, N))) return bail(st);
  if ((st = halloc(h, &a.cellrec, 4 * (size_t)N))) return bail(st);
  if ((st = halloc(h, &a.rec_b, 4 * (size_t)N))) return bail(st);
  if ((st = halloc(h, &a.dbg_a, N)) || (st = halloc(h, &a.dbg_b, N))) return bail(st);
  if ((st = halloc(h, &a.sorted, a.K))) return bail(st);
  if ((st = halloc(h, &a.work, 2 * (size_t)a.K))) return bail(st);
  if ((st = halloc(h, &a.act_flags, 2 * (size_t)a.K))) return bail(st);
  a.huge_slots = huge_slots(N, a.K);
  if ((st = halloc(h, &a.huge, (size_t)a.huge_slots * 4 * kMaxNpad))) return bail(st);
  {
    int32_t *d_cm, *d_rm;
    if ((st = upload(h, &d_cm, hp.colmap.data(), hp.colmap.size()))) return bail(st);
    if ((st = upload(h, &d_rm, hp.rowmap.data(), hp.rowmap.size()))) return bail(st);
    a.colmap = d_cm;
    a.rowmap = d_rm;
    int32_t *d_xs, *d_ys, *d_xs0, *d_ys0;
    if ((st = upload(h, &d_xs, hp.xseg.data(), hp.xseg.size()))) return bail(st);
    if ((st = upload(h, &d_ys, hp.yseg.data(), hp.yseg.size()))) return bail(st);
    if ((st = upload(h, &d_xs0, hp.xseg0.data(), hp.xseg0.size()))) return bail(st);
    if ((st = upload(h, &d_ys0, hp.yseg0.data(), hp.yseg0.size()))) return bail(st);
    a.xseg = d_xs; a.yseg = d_ys; a.xseg0 = d_xs0; a.yseg0 = d_ys0;
    a.nxs = (int)hp.xseg0.size();
    a.nys = (int)hp.yseg0.size();
    if ((st = halloc(h, &a.seg_stamp, (size_t)a.nxs * a.nys))) return bail(st);
    if ((st = halloc(h, &a.m0_hist, std::max(p->max_admm_iter, 1)))) return bail(st);
    // Lazy duals for a server's batched iterations, whose dual pass competes
    // with the partition kernels of every resident request (served slice of
    // 200 requests: 5.17 -> 4.71 s).  A solo handle keeps the eager pass:
    // inline on small grids it costs less than the lazy staging adds (92
    // against 100 us per iteration under 5k cells), and on large grids it
    // hides behind the partition kernels on the side stream (mixed_1m: 149
    // against 165 us per partition launch).  The adaptive search reads
    // prices anywhere in the grid, so it keeps every price current.
    a.lazy = !p->use_adaptive_search && pool != nullptr ? 1 : 0;
  }
  if (hp.tech) {
    a.tech = 1;
    int8_t* d_pm;
    int32_t *d_ce, *d_cm = nullptr, *d_mp = nullptr, *d_pins = nullptr;
    if ((st = upload(h, &d_pm, hp.pmetal.data(), P))) return bail(st);
    if ((st = upload(h, &d_ce, hp.cedge.data(), N))) return bail(st);
    if (!hp.cmacro.empty()) {
      if ((st = upload(h, &d_cm, hp.cmacro.data(), N))) return bail(st);
      if ((st = upload(h, &d_mp, hp.mpin.data(), hp.mpin.size()))) return bail(st);
      if ((st = upload(h, &d_pins, hp.pins.data(), hp.pins.size()))) return bail(st);
    }
    a.pmetal = d_pm;
    a.cedge = d_ce;
    a.cmacro = d_cm;
    a.mpin = d_mp;
    a.pins = reinterpret_cast<const int4*>(d_pins);
    for (int i = 0; i < 5; i++) a.rules[i] = hp.rules[i];
    a.tech_penalty = p->tech_penalty;
    a.edge_pen = p->edge_spacing_penalty;
    a.rhg = hp.row_height / hp.site_width;  // row_height_in_grid (AdmmKernels.hpp:701)
    if ((st = halloc(h, &a.edgeL, P)) || (st = halloc(h, &a.edgeR, P)) || (st = halloc(h, &a.edgeLs, P)) ||
        (st = halloc(h, &a.edgeRs, P)))
      return bail(st);
  }
  if ((st = halloc(h, &a.ovf_flag, N))) return bail(st);
  if ((st = halloc(h, &a.ovf_list, N))) return bail(st);
  if ((st = halloc(h, &a.trace, std::max(p->max_admm_iter, 1)))) return bail(st);
  if (ext_ctl) {
    a.ctl = ext_ctl;
  } else if ((st = halloc(h, &a.ctl, 1))) {
    return bail(st);
  }

  a.max_iter = p->max_admm_iter;
  a.terminate_thr = p->admm_terminate_thr;
  a.iter_update_partition = p->iter_update_partition;
  a.iter_threshold_dual_update = p->iter_threshold_dual_update;
  a.iter_update_dual_step = p->iter_update_dual_step;
  a.iter_perturbation_on = p->iter_perturbation_on;
  a.max_disp_in_row_height = p->max_disp_in_row_height;
  a.flag_iccad17 = p->flag_iccad17;
  a.standard_admm = p->standard_admm;
  a.max_disp_coeff = p->max_disp_coeff;
  a.coeff_update_dual_step = p->coeff_to_update_dual_step;
  a.sort_policy = p->sorting_policy;
  a.sort_algo = p->sorting_algorithm;
  a.use_adaptive = p->use_adaptive_search;
  a.stagnation_window = p->stagnation_window;
  a.min_iter_adaptive = p->min_iter_to_call_adaptive_search;
  a.cooltime_adaptive = p->cooltime_to_call_adaptive_search;
  a.ovf_ratio_adaptive = p->ovf_ratio_to_call_adaptive_search;
  a.coeff_neglect_disp = p->coeff_to_neglect_displace;
  // a lower limit (tests of the ranked sort keys) may come from the
  // environment; keys of up to 22 bits pack with a 10-bit slot
  a.pack_limit = (1 << 22) - 1;
  if (const char* e = std::getenv("SKYLINE_LG_PACK_LIMIT")) {
    const long v = std::strtol(e, nullptr, 10);
    if (v >= 0 && v < a.pack_limit) a.pack_limit = (int)v;
  }
  *out = h;
  return LG_OK;
}


// lg_solve up to the first iteration (CudaAdmmLegalizer::solve,
// CudaAdmmLegalizer.cu:250-281): control block, positions given at
// lg_create, usage, rho / lambda0, cell records, initial duals and the
// activity flags of iteration 0.  Asynchronous on `s`.
int solve_begin(LgHandle* h, hipStream_t s) {
  Args& a = h->a;
  const int N = a.N;
  const int nb = (N + 255) / 256;
  Ctl init;
  std::memset(&init, 0, sizeof(init));
  init.min_ovf = INT_MAX;
  init.last_slow = -1;
  init.npad = 1;
  HIP_TRY(hipMemcpyAsync(a.ctl, &init, sizeof(Ctl), hipMemcpyHostToDevice, s));
  // every solve starts from the positions given at lg_create
  if (N > 0) {
    HIP_TRY(hipMemcpyAsync(a.lx, h->init_lx, N * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(a.ly, h->init_ly, N * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(a.need_rot, 0, N * sizeof(int32_t), s));
  }
  h->posted = 0;
  h->solved = 0;
  h->host_ctl.iter = 0;
  HIP_TRY(hipMemsetAsync(a.usage, 0, h->P * sizeof(int32_t), s));
  if (a.tech) {  // updatePixelEdgeMapInGrid (CudaAdmmLegalizer.cu:270), live and snapshot
    hipLaunchKernelGGL(k_edge_clear, dim3(2048), dim3(256), 0, s, a);
    if (N > 0) hipLaunchKernelGGL(k_edge_init, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_edge_snap, dim3(2048), dim3(256), 0, s, a);
  }
  if (N > 0) {
    hipLaunchKernelGGL(k_usage_scatter, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_demand, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_rho, dim3(1), dim3(1), 0, s, a, h->coeff_init_rho);
    hipLaunchKernelGGL(k_inv_aug, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_lambda0, dim3(1), dim3(1), 0, s, a, h->coeff_init_dual);
    hipLaunchKernelGGL(k_pack_rec, dim3(nb), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(k_fill_dual, dim3(2048), dim3(256), 0, s, a);
  const dim3 pix_grid((a.x_size + 1023) / 1024, a.y_size);  // four pixels per thread
  HIP_TRY(hipMemsetAsync(a.act_flags, 0, 2 * (size_t)a.K * sizeof(int32_t), s));
  // every price holds the state before iteration 0
  HIP_TRY(hipMemsetAsync(a.seg_stamp, 0xff, (size_t)a.nxs * a.nys * sizeof(int32_t), s));
  HIP_TRY(hipMemsetAsync(a.bcnt, 0, std::max(a.K, 1) * sizeof(int32_t), s));
#ifdef LG_PHASE_TIMERS
  if (N > 0) {  // per-iteration block-time maxima / sums (timed builds)
    HIP_TRY(hipMemsetAsync(a.dbg_a, 0, N * sizeof(int32_t), s));
    HIP_TRY(hipMemsetAsync(a.dbg_b, 0, N * sizeof(int32_t), s));
  }
#endif
  hipLaunchKernelGGL(k_classify_init, pix_grid, dim3(256), 0, s, a);
  HIP_TRY(hipGetLastError());
  return LG_OK;
}

// One ADMM iteration of one handle (updatePartition, executeSorting, primal
// and dual updates, examineIteration); kernels read the iteration index and
// the done flag from the device-resident Ctl.
// With lazy duals the dual pass (the catch-up pass then) runs only when
// `catchup` is set; otherwise the prices of partitions without a primal step
// wait for their next staging.
int launch_iteration(LgHandle* h, hipStream_t s, bool catchup) {
  Args& a = h->a;
  const int nb = (a.N + 255) / 256;
  const dim3 pix_grid((a.x_size + 1023) / 1024, a.y_size);
  const bool pass = !a.lazy || catchup;
  const int nseg_blocks = (a.nxs * a.nys + 255) / 256;
  // cells of the listed partitions into their buckets
  if (a.N > 0) hipLaunchKernelGGL(k_bucket, dim3((a.N + kBucketCells - 1) / kBucketCells), dim3(256), 0, s, a);
  if (a.K > 8 * 1024) hipLaunchKernelGGL((k_scan<true, 1024>), dim3(1), dim3(1024), 0, s, a);
  else if (a.K > 8 * 256) hipLaunchKernelGGL((k_scan<false, 1024>), dim3(1), dim3(1024), 0, s, a);
  else hipLaunchKernelGGL((k_scan<false, 256>), dim3(1), dim3(256), 0, s, a);
  // the dual pass needs the step k_scan sets; the pixels it updates belong
  // to partitions the primal step leaves alone, so it runs on the side
  // stream beside the partition kernels, joined before k_finish
  const bool fork = h->side != nullptr && pass;
  if (fork) {
    HIP_TRY(hipEventRecord(h->ev_fork, s));
    HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    hipLaunchKernelGGL(k_dual_pass, pix_grid, dim3(256), 0, h->side, a);
    if (a.lazy) hipLaunchKernelGGL(k_stamp_inactive, dim3(nseg_blocks), dim3(256), 0, h->side, a);
    HIP_TRY(hipEventRecord(h->ev_join, h->side));
  }
  if (a.use_adaptive) {
    hipLaunchKernelGGL(k_extract_congested, dim3(nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_adaptive, dim3(1), dim3(1024), 0, s, a);
  }
  int st;
  if ((st = launch_partition<MODE_PRIMAL>(h, s))) return st;
  // the edge maps as the next step (or the final refine) begins
  if (a.tech) hipLaunchKernelGGL(k_edge_snap, dim3(2048), dim3(256), 0, s, a);
  if (fork) {
    HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
  } else if (pass) {
    hipLaunchKernelGGL(k_dual_pass, pix_grid, dim3(256), 0, s, a);
    if (a.lazy) hipLaunchKernelGGL(k_stamp_inactive, dim3(nseg_blocks), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(1), 0, s, a);
  return LG_OK;
}

// side stream and fork / join events of the dual pass
int ensure_side(LgHandle* h) {
  if (h->side) return LG_OK;
  HIP_TRY(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  HIP_TRY(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
  return LG_OK;
}

// kGraphIters iterations captured into one graph (once per handle: the
// launches depend only on the handle's fixed arguments).  The capture only
// records; nothing runs until the graph is launched.
int capture_iterations(LgHandle* h) {
  if (h->gexec) return LG_OK;
  if (!h->cap) HIP_TRY(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
  int st = LG_OK;
  // with lazy duals the graph's last iteration carries the catch-up pass
  for (int i = 0; i < kGraphIters && !st; i++) st = launch_iteration(h, h->cap, i == kGraphIters - 1);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->cap, &g);
  if (st || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st ? st : fail(LG_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    h->gexec = nullptr;
    return fail(LG_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  return LG_OK;
}

// After the last iteration, with h->host_ctl holding the device Ctl: limit
// errors, the final refine of a converged solve (doFinalRefine,
// CudaAdmmLegalizer.cu:1011-1115), positions back to the SoA arrays and the
// overflow trace.  Synchronises `s`.
int solve_end(LgHandle* h, hipStream_t s) {
  Args& a = h->a;
  const int nb = (a.N + 255) / 256;
  int st;
  if (a.N > 0 && h->p.max_admm_iter > 0) {
    if (h->host_ctl.error == kErrPartitionTooDense)
      return fail(LG_ERR_STATE, "more than 8192 cells in one partition");
    if (h->host_ctl.error == kErrUsageRange)
      return fail(LG_ERR_STATE, "pixel usage above 65535");
    if (a.lazy) {  // every price up to the last iteration, before the refine moves cells
      const dim3 g1((a.x_size + 255) / 256, a.y_size);
      hipLaunchKernelGGL(k_dual_final, g1, dim3(256), 0, s, a);
    }
    if (h->host_ctl.converged) {
      hipLaunchKernelGGL(k_refine_bucket, dim3(nb), dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_refine_scan, dim3(1), dim3(1024), 0, s, a);
      if ((st = launch_partition<MODE_REFINE>(h, s))) return st;
    }
    hipLaunchKernelGGL(k_unpack_pos, dim3(nb), dim3(256), 0, s, a);
    if ((st = read_ctl(h, s))) return st;
  }
  h->trace_n = std::min(h->host_ctl.iter, h->p.max_admm_iter);
  h->trace.resize(std::max(h->trace_n, 1));
  if (h->trace_n > 0)
    HIP_TRY(hipMemcpy(h->trace.data(), a.trace, h->trace_n * sizeof(int32_t), hipMemcpyDeviceToHost));
  h->solved = 1;
  h->checked = 0;
  return LG_OK;
}

}  // namespace

extern "C" {

const char* lg_last_error(void) { return g_err.c_str(); }

int lg_create(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, LgHandle** out) {
  if (!out) return fail(LG_ERR_ARG, "null argument");
  HostPrep hp;
  const auto t0 = std::chrono::steady_clock::now();
  int st = prepare_host(g, c, p, hp);
  if (st) return st;
  const auto t1 = std::chrono::steady_clock::now();
  st = create_from_prep(hp, nullptr, nullptr, nullptr, out);
  if (!st && p->verbose) {
    const auto t2 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[lg] create: host layout %.3f s, device buffers and upload %.3f s\n",
                 std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  }
  return st;
}

}  // extern "C"

int lg_create_device_grid(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, void* d_pcode,
                          LgHandle** out) {
  if (!out || !d_pcode) {
    (void)hipFree(d_pcode);
    return fail(LG_ERR_ARG, "null argument");
  }
  HostPrep hp;
  int st = prepare_host(g, c, p, hp, true);
  if (st) {
    (void)hipFree(d_pcode);
    return st;
  }
  hp.dev_pcode = d_pcode;
  // create_from_prep owns the buffer from its first allocation on; before it
  // a failure frees it here
  return create_from_prep(hp, nullptr, nullptr, nullptr, out);
}

extern "C" {

int lg_solve(LgHandle* h, void* stream) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  hipStream_t s = pick(h, stream);
  Args& a = h->a;
  int st;
  if ((st = solve_begin(h, s))) return st;
  if (a.N > 0 && h->p.max_admm_iter > 0) {
    // the dual pass runs beside the partition kernels only on large grids:
    // below kSideMinPixels the fork / join events cost more than the overlap
    // saves (measured 92 against 102 us per iteration under 5k cells, 128
    // against 129 at 80k-200k cells; the 19.2M-pixel 1M-cell grid gains)
    if (h->P >= kSideMinPixels && (st = ensure_side(h))) return st;
    const bool graph = h->graphs && !h->prof;
    if (graph && (st = capture_iterations(h))) return st;
    int batch = kGraphIters;  // then 16, 32, 64 iterations between checks
    for (;;) {
      if (graph) {
        for (int b = 0; b < batch; b += kGraphIters) HIP_TRY(hipGraphLaunch(h->gexec, s));
      } else {
        for (int b = 0; b < batch; b++) {
          const int it = h->host_ctl.iter + b;  // the device's iteration, unless it finishes first
          if ((st = launch_iteration(h, s, it % kCatchupPeriod == kCatchupPeriod - 1))) return st;
        }
      }
      HIP_TRY(hipGetLastError());
      if ((st = read_ctl(h, s))) return st;
      if (h->host_ctl.error) break;
      if (h->host_ctl.done) break;
      batch = std::min(batch * 2, 64);
    }
  } else {
    HIP_TRY(hipMemsetAsync(a.ctl, 0, sizeof(int32_t) * 2, s));
    if ((st = read_ctl(h, s))) return st;
  }
  return solve_end(h, s);
}

int lg_postprocess(LgHandle* h, void* stream) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  if (!h->solved) return fail(LG_ERR_STATE, "lg_postprocess before lg_solve");
  hipStream_t s = pick(h, stream);
  const int nb = (h->a.N + 255) / 256;
  if (h->a.N > 0) hipLaunchKernelGGL(k_rotate, dim3(nb), dim3(256), 0, s, h->a);
  HIP_TRY(hipGetLastError());
  if (h->p.qp_admm_on) {
    int st = lg_qp_refine(h, s);
    if (st) return st;
  }
  HIP_TRY(hipStreamSynchronize(s));
  h->posted = 1;
  h->checked = 0;
  return LG_OK;
}

int lg_check(LgHandle* h, LgVioStat* out) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  hipStream_t s = h->own_stream;
  Args& a = h->a;
  const int nb = (a.N + 255) / 256;
  // CudaDatabase::checkPlace (CudaDatabase.cu:326-353)
  HIP_TRY(hipMemsetAsync(a.usage, 0, h->P * sizeof(int32_t), s));
  // disp_sum, disp_max, check_ovf, row_vio, the four strict counts and the
  // three technology counts
  static_assert(offsetof(Ctl, edge_vio) - offsetof(Ctl, disp_sum) == 10 * sizeof(unsigned long long),
                "check counters must be contiguous");
  HIP_TRY(hipMemsetAsync(&a.ctl->disp_sum, 0, 11 * sizeof(unsigned long long), s));
  if (a.N > 0) hipLaunchKernelGGL(k_usage_scatter, dim3(nb), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_overflow, dim3(2048), dim3(256), 0, s, a);
  if (a.N > 0) hipLaunchKernelGGL(k_cell_stats, dim3(nb), dim3(256), 0, s, a);
  if (a.tech) {
    // countPinVio / countEdgeSpacingVio (CudaDatabase.cu:341-350) on edge
    // maps rebuilt from the final positions (the reference counts on the
    // maps its solve left, which the QP refine's x moves do not update)
    hipLaunchKernelGGL(k_edge_clear, dim3(2048), dim3(256), 0, s, a);
    if (a.N > 0) {
      hipLaunchKernelGGL(k_edge_init, dim3(nb), dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_tech_check, dim3(nb), dim3(256), 0, s, a);
    }
  }
  HIP_TRY(hipGetLastError());
  int st = read_ctl(h, s);
  if (st) return st;
  const Ctl& k = h->host_ctl;
  h->vio = LgVioStat{(int32_t)k.check_ovf, (int32_t)k.row_vio, (int32_t)k.pin_short, (int32_t)k.pin_access,
                     (int32_t)(k.edge_vio / 2), (int32_t)k.site_vio, (int32_t)k.fence_vio, (int32_t)k.core_vio,
                     (int32_t)k.rail_vio};
  h->checked = 1;
  if (out) *out = h->vio;
  return LG_OK;
}

int lg_copy_out(LgHandle* h, int32_t* lx, int32_t* ly, int32_t* rot) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(h->own_stream));
  size_t n = (size_t)h->a.N * sizeof(int32_t);
  if (n == 0) return LG_OK;
  if (lx) HIP_TRY(hipMemcpy(lx, h->a.lx, n, hipMemcpyDeviceToHost));
  if (ly) HIP_TRY(hipMemcpy(ly, h->a.ly, n, hipMemcpyDeviceToHost));
  if (rot) HIP_TRY(hipMemcpy(rot, h->a.need_rot, n, hipMemcpyDeviceToHost));
  return LG_OK;
}

int lg_get_stats(LgHandle* h, LgStats* out) {
  if (!h || !out) return fail(LG_ERR_ARG, "null argument");
  if (!h->checked) {
    int st = lg_check(h, nullptr);
    if (st) return st;
  }
  std::memset(out, 0, sizeof(*out));
  const Ctl& c = h->host_ctl;
  out->iterations = h->trace_n;
  out->converged = c.converged;
  out->final_overflow = (int64_t)c.last_ovf;
  out->total_disp_dbu = (int64_t)c.disp_sum;
  out->max_disp_dbu = (int64_t)c.disp_max;
  out->rho = c.rho;
  out->initial_dual = c.lambda0;
  out->num_cells = h->a.N;
  out->num_adaptive_calls = c.adaptive_calls;
  out->vio = h->vio;
  return LG_OK;
}

int lg_get_pixel_state(LgHandle* h, int32_t* usage, float* dual) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(h->own_stream));
  // device arrays are row-major; the ABI returns x-major
  const int XS = h->a.x_size, YS = h->a.y_size;
  std::vector<int32_t> tmp(h->P);
  if (usage) {
    HIP_TRY(hipMemcpy(tmp.data(), h->a.usage, h->P * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int y = 0; y < YS; y++)
      for (int x = 0; x < XS; x++) usage[(size_t)x * YS + y] = tmp[(size_t)y * XS + x];
  }
  if (dual) {
    HIP_TRY(hipMemcpy(tmp.data(), h->a.dual, h->P * sizeof(float), hipMemcpyDeviceToHost));
    for (int y = 0; y < YS; y++)
      for (int x = 0; x < XS; x++) std::memcpy(&dual[(size_t)x * YS + y], &tmp[(size_t)y * XS + x], 4);
  }
  return LG_OK;
}

int lg_get_ovf_trace(LgHandle* h, int32_t* trace, int32_t cap, int32_t* n) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  int m = std::min(cap, (int32_t)h->trace_n);
  if (trace && m > 0) std::memcpy(trace, h->trace.data(), m * sizeof(int32_t));
  if (n) *n = h->trace_n;
  return LG_OK;
}

// Debug counters of a LG_PHASE_TIMERS build (zeros otherwise).
int lg_debug_phase_ticks(LgHandle* h, uint64_t* out, int n) {
  if (!h || !out) return fail(LG_ERR_ARG, "null argument");
  Ctl c;
  HIP_TRY(hipMemcpy(&c, h->a.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  for (int i = 0; i < n && i < 16; i++) out[i] = c.phase_ticks[i];
  return LG_OK;
}

int lg_profile_partition(LgHandle* h, int enable) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  if (enable && h->ev.empty()) {
    h->ev.resize(3 * 64);
    for (auto& e : h->ev) HIP_TRY(hipEventCreate(&e));
  }
  h->prof = enable ? 1 : 0;
  h->ev_used = 0;
  h->prof_ms[0] = h->prof_ms[1] = 0;
  h->prof_n[0] = h->prof_n[1] = 0;
  return LG_OK;
}

int lg_use_graphs(LgHandle* h, int32_t enable) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  h->graphs = enable ? 1 : 0;
  return LG_OK;
}

int lg_profile_read(LgHandle* h, double* ms, int64_t* launches) {
  if (!h) return fail(LG_ERR_ARG, "null handle");
  for (int c = 0; c < 2; c++) {
    if (ms) ms[c] = h->prof_ms[c];
    if (launches) launches[c] = h->prof_n[c];
  }
  return LG_OK;
}

#ifdef LG_PHASE_TIMERS
int lg_debug_iter_times(LgHandle* h, int32_t* maxt, int32_t* sumt, int32_t* slow, int n) {
  n = std::min(n, h->a.N / 16);
  HIP_TRY(hipMemcpy(maxt, h->a.dbg_a, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(sumt, h->a.dbg_b, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(slow, h->a.ovf_list, 16 * n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return LG_OK;
}
#endif

int lg_destroy(LgHandle* h) {
  if (!h) return LG_OK;
  for (auto& e : h->ev) (void)hipEventDestroy(e);
  if (h->side) {
    (void)hipStreamSynchronize(h->side);
    (void)hipStreamDestroy(h->side);
  }
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  for (void* b : h->bufs) {
    if (h->pool) h->pool->release(b);
    else (void)hipFree(b);
  }
  if (h->pool && h->blk) h->pool->release(h->blk);
  if (h->own_stream && h->owns_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return LG_OK;
}

}  // extern "C"

// QP refine (post-process), same translation unit
#include "lg_qp.inc.hip"

// request server, same translation unit
#include "lg_server.inc.hip"
