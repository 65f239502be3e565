// HIP kernels of the ADMM negotiation legalizer for gfx950 (MI355X).
//
// Data layout in HBM (one legalization handle):
//   pixels, row-major by grid row (pixel_id = y * x_size + x), so a partition
//   row of x_hint sites is one contiguous run:
//     usage  int32   cells covering the pixel
//     pcode  int16   valid ? fence_group + 2 : 0   (0 = not placeable)
//     dual   float   negotiation price (Lagrange multiplier)
//     region int32   region id (adaptive search projection only)
//   cells, SoA int32: lx, ly (grid), w, h, code (= group + 2), gnd, ox, oy (dbu)
//
// One iteration = k_bucket (cells of the listed partitions into per-partition
// buckets) -> k_scan -> k_partition<PRIMAL> (sort + sequential cell moves + fused dual
// update / overflow / next activity flags) -> k_dual_pass -> k_finish.
// Every kernel reads the iteration index and the done flag from Ctl in
// device memory, so batches of iterations are launched without host syncs.
//
// Floating point follows one contract with the CPU oracle: IEEE binary32,
// no contraction (built with -ffp-contract=off), IEEE division, and the
// reference's double-promoted quadratic term in the dual update.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "lg_common.h"

namespace lgdev {

// Candidate moves: Legalizer::makeDirectionVectors (Legalizer.cpp:249-354).
__constant__ int2 c_dir_default[kNumDirDefault];
__constant__ int2 c_dir_wide[kNumDirWide];

constexpr float kInf = __builtin_huge_valf();

// Optional per-phase timing of the partition kernel (build with
// -DLG_PHASE_TIMERS): lane 0 of every workgroup adds the wall-clock ticks of
// each phase to Ctl::phase_ticks.
#ifdef LG_PHASE_TIMERS
#define PHASE_MARK(i)                                                          \
  do {                                                                         \
    unsigned long long t_ = wall_clock64();                                    \
    if (threadIdx.x == 0) atomicAdd(&a.ctl->phase_ticks[i], t_ - t_phase_);    \
    ph_[(i) & 7] = (int)(t_ - t_phase_);                                       \
    t_phase_ = t_;                                                             \
  } while (0)
#define PHASE_START()                                                          \
  unsigned long long t_phase_ = wall_clock64(), t_begin_ = t_phase_;           \
  int ph_[8] = {0, 0, 0, 0, 0, 0, 0, 0}
// per-iteration slowest block: total and phase times in ovf_list[8 * iter ..]
#define PHASE_END()                                                            \
  do {                                                                         \
    unsigned long long d_ = wall_clock64() - t_begin_;                         \
    if (threadIdx.x == 0) {                                                    \
      atomicAdd(&a.ctl->phase_ticks[9], d_);                                   \
      atomicMax(&a.ctl->phase_ticks[10], d_);                                  \
      if (MODE == MODE_PRIMAL && iter < a.N) {                                 \
        int old_ = atomicMax(&a.dbg_a[iter], (int)d_);                         \
        atomicAdd(&a.dbg_b[iter], (int)d_);                                    \
        if ((int)d_ > old_ && 16 * iter + 15 < a.N) {                          \
          int* o_ = a.ovf_list + 16 * iter;                                    \
          o_[0] = (int)d_; o_[1] = ph_[0]; o_[2] = ph_[1]; o_[3] = ph_[2];     \
          o_[4] = ph_[3]; o_[5] = count; o_[6] = (int)sub_acc_[5]; o_[7] = k;  \
          o_[8] = ph_[5]; o_[9] = ph_[6]; o_[10] = ph_[7];                     \
          for (int q_ = 0; q_ < 5; q_++) o_[11 + q_] = (int)sub_acc_[q_];      \
        }                                                                      \
      }                                                                        \
    }                                                                          \
  } while (0)
#define PHASE_COUNT(i) \
  do { if (threadIdx.x == 0) atomicAdd(&a.ctl->phase_ticks[i], 1ull); } while (0)
#define SUB_START() unsigned long long t_sub_ = wall_clock64()
#define SUB_MARK(i)                                                            \
  do {                                                                         \
    unsigned long long t_ = wall_clock64();                                    \
    sub_acc_[(i) - 11] += t_ - t_sub_;                                         \
    t_sub_ = t_;                                                               \
  } while (0)
#define SUB_DECL() unsigned long long sub_acc_[6] = {0, 0, 0, 0, 0, 0}
#define SUB_FLUSH()                                                            \
  do {                                                                         \
    if (threadIdx.x == 0)                                                      \
      for (int q_ = 0; q_ < 6; q_++)                                           \
        if (sub_acc_[q_]) atomicAdd(&a.ctl->phase_ticks[q_ < 5 ? 11 + q_ : 7], sub_acc_[q_]); \
  } while (0)
#else
#define SUB_DECL() do {} while (0)
#define SUB_FLUSH() do {} while (0)
#define SUB_START() do {} while (0)
#define SUB_MARK(i) do {} while (0)
#define PHASE_END() do {} while (0)
#define PHASE_COUNT(i) do {} while (0)
#define PHASE_MARK(i) do {} while (0)
#define PHASE_START() do {} while (0)
#endif

// std::max(a, b) of the reference's host-side semantics: (a < b) ? b : a
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

// delta of dual_update_functor_with_step_size for usage ui: the literal 0.5
// promotes the quadratic term to double before the float rounding
__device__ __forceinline__ float dual_delta(int ui, float rho, float neg) {
  float u = float(ui);
  float gg = fmax_ref(u - 1.0f, neg);
  double q = (double)rho * 0.5;
  q = q * (double)gg;
  q = q * (double)gg;
  return (float)((double)gg + q);
}

struct Args {
  // grid
  int x_size, y_size, site_width, row_height, core_lx, core_ly;
  int x_hint, y_hint, ncol, nrow, K;
  const int16_t* pcode;
  const int32_t* pregion;
  const int32_t* vdd_up;
  int32_t* usage;
  float* dual;
  // regions (shrunk grid bounds)
  const int32_t* rxmin;
  const int32_t* rymin;
  const int32_t* rxmax;
  const int32_t* rymax;
  // cells
  int N;
  int32_t* lx;
  int32_t* ly;
  const int32_t* w;
  const int32_t* h;
  const int32_t* code;
  const int32_t* gnd;
  const int32_t* ox;
  const int32_t* oy;
  const int32_t* rand_prio;
  const int32_t* k_height_num;
  int32_t* need_rot;
  // partitioning
  int32_t* bucket;     // [K * kMaxNpad] cells of each listed partition, in arrival order
  int32_t* bcnt;       // [K] cells appended to each bucket this iteration (zeroed by k_scan)
  int32_t* pcnt;       // [K] cells the partition kernel processes per partition
  int32_t* off;        // [K + 2] exclusive scan of pcnt: where each partition's order is kept
  int32_t* scratch;    // [K + 1] scan scratch of the final refine's grouping
  int32_t* grouped;    // [N] cell ids in processing order, partition by partition
  int32_t* trace;
  int32_t* cellrec;    // [4N] cell records (int4), the positions' home during a solve
  int32_t* rec_b;      // [4N] cell records in processing order (int4)
  int32_t *dbg_a, *dbg_b;  // [N] per-iteration timing of LG_PHASE_TIMERS builds
  int32_t* sorted;     // [K]  1 if the partition's order was sorted in the last primal step
  int32_t* work;       // [2K] partitions the primal step processes: class 0 (at most
                       //      kNpadSmall cells) from 0, class 1 from K; counts in Ctl::n_work
  int32_t* act_flags;  // [2][K] partition has a pixel with usage > 1 (by iteration parity)
  int32_t* huge;       // [huge_slots][4 * kMaxNpad] sort buffers of partitions above kMaxNpadLds slots
  int huge_slots;
  // lazy dual updates: the prices of a partition that takes no primal step
  // are brought up to date when it is next staged (or by the periodic
  // catch-up pass), from the last iteration that updated its segment
  const int32_t* xseg;  // [x_size] dual segment column of x
  const int32_t* yseg;  // [y_size] dual segment row of y
  const int32_t* xseg0; // [nxs] first x of a segment column
  const int32_t* yseg0; // [nys] first y of a segment row
  int nxs, nys;
  int32_t* seg_stamp;   // [nxs * nys] last iteration whose dual update the segment's prices hold
  float* m0_hist;       // [max_iter] what iteration k's dual update adds to a price of usage 0
  int lazy;             // 1: lazy dual updates (0 with the adaptive search)
  const int32_t* colmap;  // [3][x_size] partition column of x for each fold offset
  const int32_t* rowmap;  // [3][y_size] partition row of y for each fold offset
  int32_t* ovf_flag;   // [N] adaptive search: cell has overflow
  int32_t* ovf_list;   // [N]
  // parameters
  int max_iter, terminate_thr, iter_update_partition;
  int iter_threshold_dual_update, iter_update_dual_step, iter_perturbation_on;
  int max_disp_in_row_height, flag_iccad17, standard_admm;
  float max_disp_coeff, coeff_update_dual_step;
  int sort_policy, sort_algo;
  int use_adaptive, stagnation_window, min_iter_adaptive, cooltime_adaptive;
  float ovf_ratio_adaptive, coeff_neglect_disp;
  int pack_limit;      // sort keys at or above it are ranked before packing ((1 << 22) - 1)
  // ICCAD 2017 technology (tech != 0; AdmmKernels.hpp:193-331): pixel maps
  // row-major.  An edge-map word counts the cells whose left (edgeL) or
  // right (edgeR) edge of type 1 (low half) / type 2 (high half) lies on the
  // pixel, each half biased by 0x8000; edgeLs / edgeRs hold the maps as they
  // stood when the current primal or refine step began, which a partition
  // reads for pixels outside its own.
  int tech;
  const int8_t* pmetal;  // [P] power-metal layer over the pixel, or -1
  uint32_t* edgeL;
  uint32_t* edgeR;
  uint32_t* edgeLs;
  uint32_t* edgeRs;
  const int32_t* cedge;  // [N] left | right << 8 edge type (0, 1, 2)
  const int32_t* cmacro; // [N] master, or null (no pins)
  const int32_t* mpin;   // [masters + 1] pin range of each master
  const int4* pins;      // {bgn, end, layer, 0}: corners flattened x * h + y in the cell
  int rules[5];          // spacing in sites between edge types a and b at a + b
  float tech_penalty, edge_pen;
  int rhg;               // row height in sites
  Ctl* ctl;
};

// linear index of pixel (x, y) in HBM; grids hold fewer than 2^30 pixels
__device__ __forceinline__ unsigned pidx(const Args& a, int x, int y) {
  return (unsigned)y * (unsigned)a.x_size + (unsigned)x;
}

__device__ __forceinline__ bool correct_row(int h, int gnd, int vdd) {  // DeviceUtil.hpp:11-22
  return !((h & 1) == 0 && gnd != vdd);
}

__device__ __forceinline__ int triple_fold(const Args& a, int iter) {
  return (iter / a.iter_update_partition) % 3;
}

// ---------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------

// computeUsageGridKernel (DatabaseKernels.hpp:69-103): footprint clipped to the grid
__global__ void k_usage_scatter(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int x0 = max(0, a.lx[c]), y0 = max(0, a.ly[c]);
  int x1 = min(a.x_size, a.lx[c] + a.w[c]), y1 = min(a.y_size, a.ly[c] + a.h[c]);
  for (int x = x0; x < x1; x++)
    for (int y = y0; y < y1; y++) atomicAdd(&a.usage[pidx(a, x, y)], 1);
}

// computeDemandKernel (AdmmKernels.hpp:44-81); summed exactly in 64-bit integers
__global__ void k_demand(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (c < a.N) {
    int x0 = a.lx[c], y0 = a.ly[c];
    bool off = false;
    long long s = 0;
    for (int x = x0; x < x0 + a.w[c] && !off; x++) {
      if (x < 0 || x >= a.x_size) { off = true; break; }
      for (int y = y0; y < y0 + a.h[c]; y++) {
        if (y < 0 || y >= a.y_size) { off = true; break; }
        s += a.usage[pidx(a, x, y)];
      }
    }
    if (!off) v = (unsigned long long)s;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(&a.ctl->demand_sum, v);
}

// initializeRho (CudaAdmmLegalizer.cu:321-350)
__global__ void k_rho(Args a, float coeff_init_rho) {
  float demand = (float)a.ctl->demand_sum;
  float t = coeff_init_rho * demand;
  a.ctl->rho = t / float(a.N);
}

// computeInverseAugmentedDemandKernel (AdmmKernels.hpp:83-120); the per-cell
// FP32 values are summed in 2^-32 fixed point (order independent).
__global__ void k_inv_aug(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (c < a.N) {
    const float rho = a.ctl->rho;
    const float expected = float(a.row_height) / float(a.site_width);
    float aug = 0.0f;
    bool off = false;
    int x0 = a.lx[c], y0 = a.ly[c];
    for (int x = x0; x < x0 + a.w[c] && !off; x++) {
      if (x < 0 || x >= a.x_size) { off = true; break; }
      for (int y = y0; y < y0 + a.h[c]; y++) {
        if (y < 0 || y >= a.y_size) { off = true; break; }
        float u = float(a.usage[pidx(a, x, y)]);
        float t = rho * 0.5f;
        t = t * u;
        t = t * u;
        float s = 1.0f + u;
        aug += s + t;
      }
    }
    if (!off) {
      float q = expected / aug;
      v = (unsigned long long)(long long)(q * 4294967296.0f);
    }
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(&a.ctl->invaug_fx, v);
}

// initializeDual (CudaAdmmLegalizer.cu:383-395)
__global__ void k_lambda0(Args a, float coeff_init_dual) {
  Ctl* ctl = a.ctl;
  float sum = (float)((double)(long long)ctl->invaug_fx * (1.0 / 4294967296.0));
  float t = coeff_init_dual * sum;
  float l0 = t / float(a.N);
  ctl->lambda0 = l0;
  ctl->step = l0;
}

__global__ void k_fill_dual(Args a) {
  size_t P = (size_t)a.x_size * a.y_size;
  float l0 = a.ctl->lambda0;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x)
    a.dual[p] = l0;
}

// ---------------------------------------------------------------------------
// per-iteration kernels
// ---------------------------------------------------------------------------

// -- grouping of the cells by partition (updatePartition,
// CudaAdmmLegalizer.cu:166-228): a cell's partition under this iteration's
// fold offset, or none if its footprint spans two (assignPartitionKernel,
// AdmmKernels.hpp:122-162).  Only the partitions the primal step processes
// need their cells: every partition in the perturbation iterations or while
// a cell sits on a wrong-parity row, otherwise those whose activity flag is
// set (a partition without a pixel above usage one moves nothing).  Their
// cells are appended to a bucket of kMaxNpad slots per partition in arbitrary
// order; the partition kernel sorts its bucket by cell id, the order of the
// reference's thrust::stable_sort_by_key.  One thread per cell.
__device__ __forceinline__ int cell_partition(const Args& a, const int* cm, const int* rm, int4 r) {
  const int x = r.x & 0xffff, y = (int)((unsigned)r.x >> 16);
  const int w = r.y & 0xfff, h = (r.y >> 12) & 0x3f;
  const int c1 = cm[x], c2 = cm[x + w - 1], r1 = rm[y], r2 = rm[y + h - 1];
  return (c1 == c2 && r1 == r2) ? c1 * a.nrow + r1 : a.K;
}

__device__ __forceinline__ void append_to_bucket(const Args& a, int p, int c) {
  const int s = atomicAdd(&a.bcnt[p], 1);
  // a count above kMaxNpad makes k_scan stop the solve with an error
  if (s < kMaxNpad) a.bucket[(size_t)p * kMaxNpad + s] = c;
}

// kBucketCells cells per workgroup of 256 threads, four per thread (cells
// bx * 1024 + q * 256 + t), each step's loads issued together; a cell's
// partition from its corners by the arithmetic of the colmap / rowmap tables
__device__ __forceinline__ void bucket_body(const Args& a, int bx) {
  const Ctl* ctl = a.ctl;
  if (ctl->done) return;
  const int iter = ctl->iter;
  // the adaptive search's iteration orders no partition (count 0)
  if (a.use_adaptive && ctl->exceptional) return;
  const bool skip_ok = iter >= a.iter_perturbation_on && ctl->wrong_row == 0;
  const int tf = triple_fold(a, iter);
  const int ox = a.x_hint / 3 * tf, oy = a.y_hint / 3 * tf;
  const int* fl = a.act_flags + (size_t)(iter & 1) * a.K;
  constexpr int Q = kBucketCells / 256;
  const int c0 = bx * kBucketCells + (int)threadIdx.x;
  int2 r[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int c = c0 + q * 256;
    r[q] = c < a.N ? reinterpret_cast<const int2*>(a.cellrec)[2 * c] : make_int2(0, 0);
  }
  int p[Q], f[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int x = r[q].x & 0xffff, y = (int)((unsigned)r[q].x >> 16);
    const int w = r[q].y & 0xfff, h = (r[q].y >> 12) & 0x3f;
    const int c1 = ((x + ox) / a.x_hint) % a.ncol, c2 = ((x + w - 1 + ox) / a.x_hint) % a.ncol;
    const int r1 = ((y + oy) / a.y_hint) % a.nrow, r2 = ((y + h - 1 + oy) / a.y_hint) % a.nrow;
    // (assignPartitionKernel, AdmmKernels.hpp:122-162: none if the footprint spans two)
    p[q] = (c0 + q * 256 < a.N && c1 == c2 && r1 == r2) ? c1 * a.nrow + r1 : a.K;
  }
#pragma unroll
  for (int q = 0; q < Q; q++) f[q] = (p[q] < a.K && skip_ok) ? fl[p[q]] : 1;
#pragma unroll
  for (int q = 0; q < Q; q++)
    if (p[q] < a.K && f[q] != 0) append_to_bucket(a, p[q], c0 + q * 256);
}
__global__ __launch_bounds__(256) void k_bucket(Args a) { bucket_body(a, blockIdx.x); }

// Exclusive scan of n ints in place by one workgroup of 1024, in chunks of
// 32768: all loads of a chunk are issued before the first use (one memory
// round trip per chunk), staged through LDS with one pad word per 32 so that
// both the coalesced and the thread-contiguous accesses are conflict-free,
// 32 contiguous values per thread, wave scans, coalesced stores.
__device__ __forceinline__ int scan_pad(int j) { return j + (j >> 5); }

__device__ __forceinline__ void block_scan_inplace(int* v, int n) {
  constexpr int CHUNK = 32768, PER = CHUNK / 1024;
  __shared__ int buf[CHUNK + CHUNK / 32];
  __shared__ int wsum[16];
  __shared__ int carry;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += CHUNK) {
    const int m = min(CHUNK, n - c0);
    int y[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) y[i] = i * 1024 + t < m ? v[c0 + i * 1024 + t] : 0;
#pragma unroll
    for (int i = 0; i < PER; i++) buf[scan_pad(i * 1024 + t)] = y[i];
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      y[i] = buf[t * (PER + 1) + i];
      s += y[i];
    }
    int incl = s;
    for (int o = 1; o < kWave; o <<= 1) {
      int z = __shfl_up(incl, o);
      if (lane >= o) incl += z;
    }
    if (lane == kWave - 1) wsum[wv] = incl;
    __syncthreads();
    int run = carry + incl - s;
    for (int q = 0; q < wv; q++) run += wsum[q];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      buf[t * (PER + 1) + i] = run;
      run += y[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++)
      if (i * 1024 + t < m) v[c0 + i * 1024 + t] = buf[scan_pad(i * 1024 + t)];
    if (t == 1023) carry = run;
    __syncthreads();
  }
}

// Totals of the scan: off[K], off[K + 1], npad from the largest count, and
// the dual step schedule.  One thread.
__device__ __forceinline__ void scan_finish(const Args& a, int K, int m, int total) {
  Ctl* ctl = a.ctl;
  a.off[K] = total;
  a.off[K + 1] = total;
  int np = 1;
  while (np < m) np <<= 1;
  ctl->max_cnt = m;
  ctl->npad = np;
  ctl->n_huge = 0;
  if (np > kMaxNpad) { ctl->error = kErrPartitionTooDense; ctl->done = 1; }
  int it = ctl->iter;
  if (it % a.iter_update_dual_step == 0 && it > a.iter_threshold_dual_update) {
    float inc = a.coeff_update_dual_step * ctl->lambda0;
    ctl->step = ctl->step + inc;
  }
  if (it < a.max_iter) {
    const float rho = ctl->rho;
    a.m0_hist[it] = ctl->step * dual_delta(0, rho, -1.0f / rho);
  }
}

// scan_body for K <= 8 * NT (NT = 1024 or 256 threads): eight partitions per
// thread (i = j * NT + t), counts kept in registers and scanned there: the
// exclusive prefix of i is the sum of the (j', wave') totals before (j, wave)
// plus the wave's exclusive prefix -- three barriers, no second pass over HBM.
template <int NT>
__device__ __forceinline__ void scan_small(const Args& a, int K, bool skip_ok, bool adaptive_now,
                                           const int32_t* cf, int* s_n, int* s_max) {
  constexpr int W = NT / 64;   // waves
  constexpr int E = 8 * W;     // (row, wave) totals
  static_assert(E <= 128, "at most 16 waves");
  __shared__ int tot[E], pre[E], s_total;
  Ctl* ctl = a.ctl;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int iter = ctl->iter;
  int cn[8], fl[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int i = j * NT + t;
    cn[j] = i < K ? a.bcnt[i] : 0;
    fl[j] = (i < K && skip_ok) ? cf[i] : 1;
  }
  int mx = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int i = j * NT + t;
    const bool in = i < K;
    const int c = cn[j];
    if (in) {
      a.pcnt[i] = c;
      a.bcnt[i] = 0;
    }
    mx = max(mx, c);
    const bool run = in && fl[j] != 0;
    if (in && !run) a.sorted[i] = 0;
    const int cls = (!adaptive_now && c > kNpadSmall) ? 1 : 0;
    // list order is free: partitions of one step are independent
    const unsigned long long m0 = __ballot(run && cls == 0), m1 = __ballot(run && cls == 1);
    int b0 = 0, b1 = 0;
    if (lane == 0) {
      if (m0) b0 = atomicAdd(&s_n[0], __popcll(m0));
      if (m1) b1 = atomicAdd(&s_n[1], __popcll(m1));
    }
    b0 = __shfl(b0, 0);
    b1 = __shfl(b1, 0);
    if (run && cls == 0) a.work[b0 + __popcll(m0 & lt)] = i;
    if (run && cls == 1) a.work[K + b1 + __popcll(m1 & lt)] = i;
  }
  // activity flags of the next iteration, set by the primal step
  int32_t* nf = a.act_flags + (size_t)((iter + 1) & 1) * K;
  for (int i = t; i < K; i += NT) nf[i] = 0;
  // inclusive wave scans of the eight rows
  int incl[8];
#pragma unroll
  for (int j = 0; j < 8; j++) incl[j] = cn[j];
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int z = __shfl_up(incl[j], o);
      if (lane >= o) incl[j] += z;
    }
  }
  if (lane == kWave - 1)
#pragma unroll
    for (int j = 0; j < 8; j++) tot[j * W + wave] = incl[j];
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
  if (lane == 0) s_max[wave] = mx;
  __syncthreads();
  if (t < 2) ctl->n_work[t] = s_n[t];
  if (t < kWave) {
    // exclusive scan of the E (row, wave) totals, two per lane
    const int v0 = 2 * t < E ? tot[2 * t] : 0, v1 = 2 * t + 1 < E ? tot[2 * t + 1] : 0;
    int x = v0 + v1;
    for (int o = 1; o < kWave; o <<= 1) {
      const int z = __shfl_up(x, o);
      if (lane >= o) x += z;
    }
    const int ex = x - v0 - v1;
    if (2 * t < E) pre[2 * t] = ex;
    if (2 * t + 1 < E) pre[2 * t + 1] = ex + v0;
    if (t == kWave - 1) s_total = x;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int i = j * NT + t;
    if (i < K) a.off[i] = pre[j * W + wave] + incl[j] - cn[j];
  }
  if (t == 0) {
    int m = 0;
    for (int q = 0; q < W; q++) m = max(m, s_max[q]);
    scan_finish(a, K, m, s_total);
  }
}

// Exclusive scan of the partition cell counts, max count -> npad, dual step
// schedule (updateDualStepSize, CudaAdmmLegalizer.cu:293-294, 687-691), and
// the work lists of the primal step: every partition except those whose
// activity flag says the step is a no-op, split by sort-size class.  A
// listed partition's workgroup then starts in the first dispatch round
// instead of behind thousands of workgroups that only find their partition
// idle.  One workgroup of 1024 threads.
// LARGE: the path for more than 8192 partitions is compiled in, with its
// 135 KB of LDS; without it a scan workgroup takes little LDS and shares a
// CU with the partition kernels' workgroups (callers pick it when every K is
// at most 8192).
template <bool LARGE, int NT>
__device__ __forceinline__ void scan_body(const Args& a) {
  static_assert(!LARGE || NT == 1024, "the large path runs on 1024 threads");
  Ctl* ctl = a.ctl;
  if (ctl->done) return;
  __shared__ int s_max[16];
  __shared__ int s_n[2];
  const int K = a.K;
  const int t = threadIdx.x, lane = t & 63;
  const int iter = ctl->iter;
  const bool adaptive_now = a.use_adaptive && ctl->exceptional;
  // the same test as the partition kernel's early exit
  const bool skip_ok = iter >= a.iter_perturbation_on && ctl->wrong_row == 0 && !adaptive_now;
  const int32_t* cf = a.act_flags + (size_t)(iter & 1) * K;
  const unsigned long long lt = (1ull << lane) - 1ull;
  if (t < 2) s_n[t] = 0;
  __syncthreads();
  if constexpr (!LARGE) {
    scan_small<NT>(a, K, skip_ok, adaptive_now, cf, s_n, s_max);
    return;
  }
  if (K <= 8 * 1024) {
    scan_small<1024>(a, K, skip_ok, adaptive_now, cf, s_n, s_max);
    return;
  }
  // cells per partition from the bucket counts (off <- counts, scanned
  // below; the counters are cleared for the next iteration); eight
  // partitions per thread with every load issued first
  int mx = 0;
  for (int s0 = 0; s0 < K; s0 += 8 * 1024) {
    int cn[8], fl[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = s0 + j * 1024 + t;
      cn[j] = i < K ? a.bcnt[i] : 0;
      fl[j] = (i < K && skip_ok) ? cf[i] : 1;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = s0 + j * 1024 + t;
      const bool in = i < K;
      const int c = cn[j];
      if (in) {
        a.off[i] = c;
        a.pcnt[i] = c;
        a.bcnt[i] = 0;
      }
      mx = max(mx, c);
      const bool run = in && fl[j] != 0;
      if (in && !run) a.sorted[i] = 0;
      const int cls = (!adaptive_now && c > kNpadSmall) ? 1 : 0;
      // list order is free: partitions of one step are independent
      const unsigned long long m0 = __ballot(run && cls == 0), m1 = __ballot(run && cls == 1);
      int b0 = 0, b1 = 0;
      if (lane == 0) {
        if (m0) b0 = atomicAdd(&s_n[0], __popcll(m0));
        if (m1) b1 = atomicAdd(&s_n[1], __popcll(m1));
      }
      b0 = __shfl(b0, 0);
      b1 = __shfl(b1, 0);
      if (run && cls == 0) a.work[b0 + __popcll(m0 & lt)] = i;
      if (run && cls == 1) a.work[K + b1 + __popcll(m1 & lt)] = i;
    }
  }
  __syncthreads();
  if (t < 2) ctl->n_work[t] = s_n[t];
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
  if ((t & 63) == 0) s_max[t >> 6] = mx;
  // activity flags of the next iteration, set by the primal step
  int32_t* nf = a.act_flags + (size_t)((iter + 1) & 1) * K;
  for (int i = t; i < K; i += 1024) nf[i] = 0;
  __syncthreads();
  const int last = K > 0 ? a.off[K - 1] : 0;
  __syncthreads();
  block_scan_inplace(a.off, K);
  if (t == 0) {
    int m = 0;
    for (int q = 0; q < 16; q++) m = max(m, s_max[q]);
    scan_finish(a, K, m, K > 0 ? a.off[K - 1] + last : 0);
  }
}
template <bool LARGE, int NT>
__global__ __launch_bounds__(NT) void k_scan(Args a) { scan_body<LARGE, NT>(a); }

// Grouping for the final refine, which visits every partition under the
// fold offset of the converged iteration.  A partition ordered by the last
// primal step keeps that order (grouped, off, pcnt as it left them); the
// cells of every other partition go to its bucket, and its order is placed
// behind all kept ones.
__global__ __launch_bounds__(256) void k_refine_bucket(Args a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  const int tf = triple_fold(a, a.ctl->conv_iter);
  const int4 r = reinterpret_cast<const int4*>(a.cellrec)[c];
  const int p = cell_partition(a, a.colmap + tf * a.x_size, a.rowmap + tf * a.y_size, r);
  if (p < a.K && a.sorted[p] == 0) append_to_bucket(a, p, c);
}

// off / pcnt of the refine's grouping and npad = pow2(largest partition).
// One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_refine_scan(Args a) {
  Ctl* ctl = a.ctl;
  __shared__ int s_max[16];
  const int K = a.K, t = threadIdx.x;
  int mx = 0;
  for (int i = t; i < K; i += 1024) {
    const int c = a.sorted[i] ? 0 : a.bcnt[i];
    a.scratch[i] = c;
    mx = max(mx, a.sorted[i] ? a.pcnt[i] : c);
  }
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
  if ((t & 63) == 0) s_max[t >> 6] = mx;
  __syncthreads();
  const int base = a.off[K];  // end of the kept orders
  block_scan_inplace(a.scratch, K);
  for (int i = t; i < K; i += 1024) {
    if (!a.sorted[i]) {
      a.off[i] = base + a.scratch[i];
      a.pcnt[i] = a.bcnt[i];
    }
    a.bcnt[i] = 0;
  }
  if (t == 0) {
    int m = 0;
    for (int q = 0; q < 16; q++) m = max(m, s_max[q]);
    int np = 1;
    while (np < m) np <<= 1;
    ctl->npad = np;
    ctl->n_huge = 0;
    if (np > kMaxNpad) ctl->error = kErrPartitionTooDense;
  }
}


// ---------------------------------------------------------------------------
// the partition kernel: one 64-lane wavefront owns one partition
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T wave_bcast0(T v) { return __shfl(v, 0); }

__device__ __forceinline__ int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sorting networks over LDS with 64 lanes; pairs of one step are disjoint.
// bitonicSortDescending (AdmmKernels.hpp:899-945), the reference's network.
__device__ void lds_bitonic_desc(int n, int* key, int* val) {
  const int lane = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < n; i += kWave) {
        int ij = i ^ j;
        if (ij > i) {
          int ki = key[i], kj = key[ij];
          bool sw = ((i & k) == 0) ? (ki < kj) : (ki > kj);
          if (sw) {
            key[i] = kj; key[ij] = ki;
            int vi = val[i]; val[i] = val[ij]; val[ij] = vi;
          }
        }
      }
      __syncthreads();
    }
}


// mergeSortDescending (SortingKernels.hpp:143-219): bottom-up merges, a tie
// takes the right run first.  key/val hold the input; bk/bv are buffers.
__device__ void lds_merge_desc(int n, int* key, int* val, int* bk, int* bv) {
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += kWave) { bk[i] = key[i]; bv[i] = val[i]; }
  __syncthreads();
  for (int w = 1; w < n; w <<= 1) {
    for (int j = lane; j < n / 2; j += kWave) {
      int left = j * 2 * w;
      int mid = min(left + w, n), right = min(left + 2 * w, n);
      int x = left, y = mid, k = left;
      while (x < mid && y < right) {
        if (key[x] > key[y]) { bk[k] = key[x]; bv[k] = val[x]; x++; }
        else { bk[k] = key[y]; bv[k] = val[y]; y++; }
        k++;
      }
      while (x < mid) { bk[k] = key[x]; bv[k] = val[x]; k++; x++; }
      while (y < right) { bk[k] = key[y]; bv[k] = val[y]; k++; y++; }
    }
    __syncthreads();
    for (int i = lane; i < n; i += kWave) { key[i] = bk[i]; val[i] = bv[i]; }
    __syncthreads();
  }
}

// Segmented radix sort of the CUB option: stable, descending.  Keys are made
// unique by the input position (ties keep ascending position).
__device__ void lds_stable_desc(int n, int count, int* key, int* val, int* pos) {
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += kWave) pos[i] = i;
  __syncthreads();
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < n; i += kWave) {
        int ij = i ^ j;
        if (ij > i) {
          // "before" = larger key, then smaller position
          bool ij_before_i = (key[ij] > key[i]) || (key[ij] == key[i] && pos[ij] < pos[i]);
          bool desc = (i & k) == 0;
          if (desc ? ij_before_i : !ij_before_i) {
            int t0 = key[i]; key[i] = key[ij]; key[ij] = t0;
            t0 = val[i]; val[i] = val[ij]; val[ij] = t0;
            t0 = pos[i]; pos[i] = pos[ij]; pos[ij] = t0;
          }
        }
      }
      __syncthreads();
    }
  (void)count;
}

// Partition geometry of one iteration.  The reference paints every pixel with
// its partition colour (paintPixelsKernel, AdmmKernels.hpp:164-191):
// colour(x, y) = ((x+ox)/x_hint % ncol) * nrow + (y+oy)/y_hint % nrow.  Here
// the colour test is two interval checks per axis: partition column `col`
// covers x in [col*xh - ox, (col+1)*xh - ox) and, through the wrap-around of
// the shifted grid, [(col+ncol)*xh - ox, (col+ncol+1)*xh - ox).
struct PartGeo {
  int oxs, oys, col, row, xh, yh, ncol, nrow, xs, ys;
  int xa0, xa1, xb0, xb1, ya0, ya1, yb0, yb1;
  __device__ void init(int xs_, int ys_) {
    xs = xs_; ys = ys_;
    xa0 = col * xh - oxs; xa1 = xa0 + xh;
    xb0 = (col + ncol) * xh - oxs; xb1 = xb0 + xh;
    ya0 = row * yh - oys; ya1 = ya0 + yh;
    yb0 = (row + nrow) * yh - oys; yb1 = yb0 + yh;
  }
  __device__ bool in_x(int x) const { return (x >= xa0 && x < xa1) || (x >= xb0 && x < xb1); }
  __device__ bool in_y(int y) const { return (y >= ya0 && y < ya1) || (y >= yb0 && y < yb1); }
  // local index (x + ox) mod xh of an in-partition (or in-grid, ncol == 1) x
  __device__ int lx(int x) const {
    int l = x - xa0;
    if (l >= xh) l -= ncol * xh;
    return l;
  }
  __device__ int ly(int y) const {
    int l = y - ya0;
    if (l >= yh) l -= nrow * yh;
    return l;
  }
  // global coordinate of local column la / row lb, or -1 if outside the grid
  __device__ int gx(int la) const {
    int x = xa0 + la;
    if (x >= 0 && x < xs) return x;
    x = xb0 + la;
    return (x >= 0 && x < xs) ? x : -1;
  }
  __device__ int gy(int lb) const {
    int y = ya0 + lb;
    if (y >= 0 && y < ys) return y;
    y = yb0 + lb;
    return (y >= 0 && y < ys) ? y : -1;
  }
};

// Cell record in HBM (int4, 16 bytes): {lx | ly << 16, w | h << 12 | gnd << 18
// | code << 19, ox, oy}.  One load brings everything a cell move needs.
__device__ __forceinline__ int rec_w(int s) { return s & 0xfff; }
__device__ __forceinline__ int rec_h(int s) { return (s >> 12) & 0x3f; }
__device__ __forceinline__ int rec_gnd(int s) { return (s >> 18) & 1; }
__device__ __forceinline__ int rec_code(int s) { return (s >> 19) & 0x1fff; }
__device__ __forceinline__ int rec_x(int p) { return p & 0xffff; }
__device__ __forceinline__ int rec_y(int p) { return (int)((unsigned)p >> 16); }
__device__ __forceinline__ int rec_pos(int x, int y) { return (x & 0xffff) | (y << 16); }

// computeDispCost (AdmmKernels.hpp:508-530); `thr` is the loop-invariant
// float(max_disp_in_row_height) * float(row_height) / float(site_width)
__device__ __forceinline__ float disp_cost(const Args& a, float disp_coeff, float thr, int nx, int ny, int ox,
                                           int oy) {
  int dx = nx * a.site_width + a.core_lx;
  int dy = ny * a.row_height + a.core_ly;
  int disp = abs(dx - ox) + abs(dy - oy);
  float ds = float(disp) / float(a.site_width);
  float ex = ds - thr;
  float p = disp_coeff * ds;
  float q = a.max_disp_coeff * fmax_ref(ex, 0.0f);
  return p + q;
}

__device__ __forceinline__ float disp_threshold(const Args& a) {
  float thr = float(a.max_disp_in_row_height) * float(a.row_height);
  return thr / float(a.site_width);
}

// getNewCellPosition (AdmmKernels.hpp:540-560)
__device__ __forceinline__ void new_pos(const Args& a, int2 d, int w, int h, int cx, int cy, int& nx, int& ny) {
  int eps = (h % 2 == 0) ? 2 : 1;
  nx = cx + d.x;
  ny = cy + eps * d.y;
  nx = max(0, min(a.x_size - w, nx));
  ny = max(0, min(a.y_size - h, ny));
}

enum { MODE_PRIMAL = 0, MODE_REFINE = 1 };

// LDS tile of the partition: three arrays with row stride rs (kRowStride
// when x_hint fits, else x_hint), element ly * rs + lx for local column lx and
// row ly:  uc = usage | code << 16,  t = penalty term,  lam = dual, where t is
// the term of computePenaltySum for the pixel's current usage u
// (AdmmKernels.hpp:380-383):
//   t = rho * max(lam * (u - 1) + lam / rho, 0)     (negotiation ADMM)
//   t = rho * max(u - 1 + lam / rho, 0)             (standard ADMM)
// or +inf on a pixel of code 0, which no cell may use.  t is recomputed,
// with the same operations, whenever u changes.
constexpr int kRowStride = 256;
struct Tile {
  uint32_t* uc;
  float* t;
  float* lam;
  int* vdd;
  uint32_t* eL;  // technology only: edge maps and power metal of the tile
  uint32_t* eR;
  int8_t* pm;
  int rs;
  __device__ int at(int lx, int ly) const { return ly * rs + lx; }
};

// -- ICCAD 2017 technology: edge maps, pin-access and edge-spacing costs ----
constexpr uint32_t kEdgeBias = 0x80008000u;  // both 16-bit counts at zero
__device__ __forceinline__ int edge_count(uint32_t v, int type) {
  return type == 1 ? (int)(v & 0xffffu) - 0x8000 : (int)(v >> 16) - 0x8000;
}
// the word to add for one edge of `type` (delta +1 / -1; type 0: none)
__device__ __forceinline__ uint32_t edge_delta(int type, int delta) {
  return type == 1 ? (uint32_t)delta : (type == 2 ? (uint32_t)delta << 16 : 0u);
}
// willBeRotated (AdmmKernels.hpp:532-538): odd heights take the row's rail
__device__ __forceinline__ bool will_rotate(int h, int vdd, int gnd) { return (h & 1) != 0 && vdd != gnd; }

// changePixelEdgeType (AdmmKernels.hpp:428-465) on the LDS tile: the left
// column of the footprint at (x0, y0) takes its left edge, the right column
// its right edge; lanes take the rows
__device__ __forceinline__ void edge_add_tile(const PartGeo& g, const Tile& T, int x0, int y0, int w, int h,
                                              int lt, int rt, int delta) {
  const int la = g.lx(x0);
  int lb = la + w - 1;
  if (lb >= g.xh) lb -= g.xh;
  const int ly0 = g.ly(y0);
  for (int i = threadIdx.x; i < h; i += blockDim.x) {
    int ly = ly0 + i;
    if (ly >= g.yh) ly -= g.yh;
    if (lt) T.eL[T.at(la, ly)] += edge_delta(lt, delta);
    if (rt) T.eR[T.at(lb, ly)] += edge_delta(rt, delta);
  }
}

// computeNumPinVio (AdmmKernels.hpp:193-264) for a footprint at local
// corner (lx0, ly0) of the tile: a pin box over power metal of its own layer
// is a short, of the layer above an access violation; each counts once
__device__ int pin_vio_tile(const Args& a, const PartGeo& g, const Tile& T, int p0, int p1, int w, int h, bool rot,
                            int lx0, int ly0) {
  int shrt = 0, acc = 0;
  for (int i = p0; i < p1; i++) {
    const int4 pn = a.pins[i];
    int bx = pn.x / h, by = pn.x - (pn.x / h) * h, ex = pn.y / h, ey = pn.y - (pn.y / h) * h;
    if (rot) {
      const int ox = bx, oy = by;
      bx = w - 1 - ex; by = h - 1 - ey;
      ex = w - 1 - ox; ey = h - 1 - oy;
    }
    const int x1 = min(bx, ex), x2 = max(bx, ex), y1 = min(by, ey), y2 = max(by, ey);
    for (int x = x1; x <= x2; x++) {
      int lx = lx0 + x;
      if (lx >= g.xh) lx -= g.xh;
      for (int y = y1; y <= y2; y++) {
        int ly = ly0 + y;
        if (ly >= g.yh) ly -= g.yh;
        const int l = T.pm[T.at(lx, ly)];
        if (l > 0) {
          if (l == pn.z) shrt = 1;
          else if (l == pn.z + 1) acc = 1;
        }
      }
    }
  }
  return shrt + acc;
}

// Edge counts of `type` in columns [x0, x1) of rows [ny, ny + h): pixels of
// the partition from the live LDS tile, others from the snapshot in HBM
// (left == true: the left-edge map, else the right-edge map)
__device__ int edge_window(const Args& a, const PartGeo& g, const Tile& T, bool left, int type, int x0, int x1,
                           int ny, int h) {
  int n = 0;
  const int xe = min(x1, a.x_size);
  for (int x = max(x0, 0); x < xe; x++) {
    if (g.in_x(x)) {
      const int lx = g.lx(x);
      int ly = g.ly(ny);
      for (int y = 0; y < h; y++) {
        n += edge_count((left ? T.eL : T.eR)[T.at(lx, ly)], type);
        if (++ly == g.yh) ly = 0;
      }
    } else {
      for (int y = ny; y < ny + h; y++) n += edge_count((left ? a.edgeLs : a.edgeRs)[pidx(a, x, y)], type);
    }
  }
  return n;
}

// computeEdgeSpacingCost (AdmmKernels.hpp:266-331): the right edges of the
// neighbours within the rule's spacing left of the cell, their left edges
// right of it, times edge_spacing_penalty
__device__ float edge_cost_tile(const Args& a, const PartGeo& g, const Tile& T, int lt, int rt, int w, int h,
                                int nx, int ny) {
  int l = 0, r = 0;
  if (lt > 0)
    l = edge_window(a, g, T, false, 1, nx - a.rules[lt + 1], nx, ny, h) +
        edge_window(a, g, T, false, 2, nx - a.rules[lt + 2], nx, ny, h);
  if (rt > 0)
    r = edge_window(a, g, T, true, 1, nx + w, nx + w + a.rules[rt + 1], ny, h) +
        edge_window(a, g, T, true, 2, nx + w, nx + w + a.rules[rt + 2], ny, h);
  const float cl = float(l), cr = float(r);
  return a.edge_pen * (cl + cr);
}

// the technology of the cell being moved
struct TechCell {
  int lt, rt;  // edge types as stored (unrotated)
  int p0, p1;  // pin range of its master
};

__device__ __forceinline__ float pen_term(int ui, float l, float rho, bool standard) {
  float u = float(ui);
  float m1 = u - 1.0f;
  float q = l / rho;
  if (standard) return rho * fmax_ref(m1 + q, 0.0f);
  float m = l * m1;
  return rho * fmax_ref(m + q, 0.0f);
}

// overflow of a footprint from the LDS tile (computeCellOverflow, AdmmKernels.hpp:393-406)
__device__ __forceinline__ int ovf_lds(const PartGeo& g, const Tile& T, int cx, int cy, int w, int h) {
  int s = 0;
  int lxi = g.lx(cx);
  const int ly0 = g.ly(cy);
  for (int x = 0; x < w; x++) {
    int lyi = ly0;
    for (int y = 0; y < h; y++) {
      s += max((int)(T.uc[T.at(lxi, lyi)] & 0xffffu) - 1, 0);
      if (++lyi == g.yh) lyi = 0;
    }
    if (++lxi == g.xh) lxi = 0;
  }
  return s;
}

// ovf_lds for a footprint that does not wrap, with row stride kRowStride and
// h <= 4: columns in groups of four, all rows at once, so sixteen LDS reads
// are in flight (the per-lane footprints of a wave differ in size)
__device__ __forceinline__ int ovf_lds_flat(const Tile& T, int lx0, int ly0, int w, int h) {
  const uint32_t* p = T.uc + ly0 * kRowStride + lx0;
  int s = 0;
  for (int x = 0; x < w; x += 4) {
    uint32_t v[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int y = 0; y < 4; y++) v[q][y] = p[y * kRowStride + x + q];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int y = 0; y < 4; y++)
        if (x + q < w && y < h) s += max((int)(v[q][y] & 0xffffu) - 1, 0);
  }
  return s;
}

__device__ __forceinline__ int ovf_tile(const PartGeo& g, const Tile& T, int cx, int cy, int w, int h) {
  if (g.ncol > 1 && g.nrow > 1 && T.rs == kRowStride && h <= 4) return ovf_lds_flat(T, g.lx(cx), g.ly(cy), w, h);
  return ovf_lds(g, T, cx, cy, w, h);
}

// the same from the pixels in HBM, for a cell inside the grid
__device__ __forceinline__ int ovf_hbm(const Args& a, int cx, int cy, int w, int h) {
  int s = 0;
  for (int x = cx; x < cx + w; x++)
    for (int y = cy; y < cy + h; y++) s += max(a.usage[pidx(a, x, y)] - 1, 0);
  return s;
}

// rip-up (delta -1) or commit (+1) of a footprint by the NW waves of the
// workgroup; refreshes the penalty terms
template <int NW>
__device__ __forceinline__ void footprint_add(const PartGeo& g, const Tile& T, int x0, int y0, int w, int h,
                                              int delta, float rho, bool standard) {
  const int n = w * h;
  const int lx0 = g.lx(x0), ly0 = g.ly(y0);
  for (int i = threadIdx.x; i < n; i += NW * kWave) {
    // (one- and two-row cells, most of them, without the division)
    const int dx = h == 1 ? i : (h == 2 ? (i >> 1) : i / h);
    int lxi = lx0 + dx, lyi = ly0 + (i - dx * h);
    if (lxi >= g.xh) lxi -= g.xh;
    if (lyi >= g.yh) lyi -= g.yh;
    const int idx = T.at(lxi, lyi);
    const uint32_t e = (uint32_t)((int)T.uc[idx] + delta);
    T.uc[idx] = e;
    T.t[idx] = (e >> 16) == 0u ? kInf : pen_term((int)(e & 0xffffu), T.lam[idx], rho, standard);
  }
}

// Penalty sums of Q consecutive footprint columns (x .. x+Q-1) of height H
// for the NC candidates of a lane, row stride kRowStride: every LDS read uses
// an immediate offset from one address per candidate.  The sums run over x,
// then y, as computePenaltySum's loops.
template <int NC, int H, int Q>
__device__ __forceinline__ void sum_cols(const float* t, const int (&base)[NC], int x, float (&s)[NC]) {
  float v[NC][Q][H];
#pragma unroll
  for (int j = 0; j < NC; j++) {
    const float* p = t + base[j] + x;
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int y = 0; y < H; y++) v[j][q][y] = p[y * kRowStride + q];
  }
#pragma unroll
  for (int j = 0; j < NC; j++)
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int y = 0; y < H; y++) s[j] += v[j][q][y];
}

template <int NC, int H>
__device__ __forceinline__ void sum_rect(const float* t, const int (&base)[NC], int w, float (&s)[NC]) {
  int x = 0;
  for (; x + 4 <= w; x += 4) sum_cols<NC, H, 4>(t, base, x, s);
  switch (w - x) {
    case 1: sum_cols<NC, H, 1>(t, base, x, s); break;
    case 2: sum_cols<NC, H, 2>(t, base, x, s); break;
    case 3: sum_cols<NC, H, 3>(t, base, x, s); break;
    default: break;
  }
}

// One step of a wave minimum of 64-bit keys hi:lo held in two registers:
// each lane takes the smaller of its key and the key DPP control CTRL brings
// it (lanes the control does not write keep their own).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void key_min_step(unsigned& hi, unsigned& lo) {
  const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, ROWMASK, 0xf, false);
  const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, ROWMASK, 0xf, false);
  const bool take = ohi < hi || (ohi == hi && olo < lo);
  hi = take ? ohi : hi;
  lo = take ? olo : lo;
}

// The NC candidates of a lane (moves d = d0 + j*64, j < NC), evaluated
// together so that their LDS reads overlap.  cost = computeDispCost + computePenaltySum
// (partition_exclusive, AdmmKernels.hpp:333-391); pin-access and edge-spacing
// terms are zero (no pins, no edge rules).  A footprint that leaves the
// partition, lands on a wrong-rail row, or touches a pixel of another group or
// an invalid pixel costs +inf; the penalty terms sum in the reference's
// x-major order.
template <int MODE, int NC>
__device__ __forceinline__ void eval_candidates(const Args& a, const PartGeo& g, const Tile& T,
                                                float disp_coeff, float thr, int cx, int cy,
                                                int w, int h, int code, int gnd, int ox, int oy, int d0,
                                                const int2 (&dir)[NC], float (&cst)[NC], int ucode,
                                                unsigned long long vmask, const TechCell& tc) {
  const bool flat = g.ncol > 1 && g.nrow > 1;  // footprints inside one interval never wrap
  // the technology terms are evaluated by the one-wave launches (NC == 4)
  const bool tech = NC == 4 && a.tech != 0;
  if (MODE == MODE_PRIMAL && !tech && ucode >= 0 && flat && T.rs == kRowStride && h <= 4 && g.yh <= 64) {
    // The common case without branches: a tile of one placement code, a
    // footprint of at most four rows.  The clamp of getNewCellPosition keeps
    // a candidate on the grid (cells fit the grid), and -- the partition's
    // two intervals per axis lying at least a partition apart -- a footprint
    // is inside the partition iff its local corner l (lx() / ly()) has
    // 0 <= l <= xh - w (resp. yh - h).  The rail of the landing row comes
    // from the mask.  Same arithmetic as the general walk below.
    const int eps = (h % 2 == 0) ? 2 : 1;
    const bool wide = w > g.xh || h > g.yh || code != ucode;
    int base[NC];
    float s4[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) {
      const int nx = max(0, min(a.x_size - w, cx + dir[j].x));
      const int ny = max(0, min(a.y_size - h, cy + eps * dir[j].y));
      cst[j] = disp_cost(a, disp_coeff, thr, nx, ny, ox, oy);
      int lx = nx - g.xa0;
      lx -= lx >= g.xh ? g.ncol * g.xh : 0;
      int ly = ny - g.ya0;
      ly -= ly >= g.yh ? g.nrow * g.yh : 0;
      const bool in = (lx >= 0) & (lx + w <= g.xh) & (ly >= 0) & (ly + h <= g.yh);
      const int rail = (int)((vmask >> (unsigned)(ly & 63)) & 1ull);
      const bool ok = in & !wide & (d0 + j * kWave < kNumDirDefault) & (((h & 1) != 0) | (rail == gnd));
      base[j] = ok ? ly * kRowStride + lx : -1;
      s4[j] = 0.0f;
    }
    int bz[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) bz[j] = base[j] < 0 ? 0 : base[j];
    switch (h) {
      case 1: sum_rect<NC, 1>(T.t, bz, w, s4); break;
      case 2: sum_rect<NC, 2>(T.t, bz, w, s4); break;
      case 3: sum_rect<NC, 3>(T.t, bz, w, s4); break;
      default: sum_rect<NC, 4>(T.t, bz, w, s4); break;
    }
#pragma unroll
    for (int j = 0; j < NC; j++) cst[j] = base[j] < 0 ? kInf : cst[j] + s4[j];
    return;
  }
  int lxi[NC], lyi0[NC];
  int bad[NC];  // 0 / 1 (ints: bool arrays get packed into bytes)
  float s[NC];
#pragma unroll
  for (int j = 0; j < NC; j++) {
    int d = d0 + j * kWave;
    bad[j] = 1;
    lxi[j] = 0;
    lyi0[j] = 0;
    s[j] = 0.0f;
    cst[j] = kInf;
    if (d < kNumDirDefault) {
      int nx, ny;
      new_pos(a, dir[j], w, h, cx, cy, nx, ny);
      cst[j] = disp_cost(a, disp_coeff, thr, nx, ny, ox, oy);
      bool ok = !(nx + w > a.x_size || ny + h > a.y_size || nx < 0 || ny < 0);
      if (ok && g.ncol > 1) ok = w <= g.xh && g.in_x(nx) && g.in_x(nx + w - 1);
      if (ok && g.nrow > 1) ok = h <= g.yh && g.in_y(ny) && g.in_y(ny + h - 1);
      if (ok) {
        lyi0[j] = g.ly(ny);
        lxi[j] = g.lx(nx);
        ok = correct_row(h, gnd, T.vdd[lyi0[j]]);
      }
      bad[j] = ok ? 0 : 1;
    }
  }
  if (ucode >= 0 && flat && T.rs == kRowStride && h <= 4) {
    // every in-grid pixel of the tile carries code `ucode`: the code test is
    // one comparison, and the walk only sums the penalty terms (a bad
    // candidate sums junk from row 0, its cost is +inf anyway)
    int base[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) {
      if (code != ucode) bad[j] = 1;
      base[j] = bad[j] ? 0 : lyi0[j] * kRowStride + lxi[j];
    }
    switch (h) {
      case 1: sum_rect<NC, 1>(T.t, base, w, s); break;
      case 2: sum_rect<NC, 2>(T.t, base, w, s); break;
      case 3: sum_rect<NC, 3>(T.t, base, w, s); break;
      default: sum_rect<NC, 4>(T.t, base, w, s); break;
    }
  } else if (flat) {
    // flat walk over the footprint in x-major order (f = x * h + y), four
    // pixels per step so that sixteen LDS reads are in flight
    int col[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) col[j] = T.at(lxi[j], lyi0[j]);
    const int n = w * h;
    int fx = 0, fy = 0;
    for (int f0 = 0; f0 < n; f0 += 4) {
      int off[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        off[q] = fy * T.rs + fx;
        if (++fy == h) { fy = 0; fx++; }
      }
      uint32_t e[4][NC];
      float tv[4][NC];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int j = 0; j < NC; j++) {
          const int i = (bad[j] || f0 + q >= n) ? 0 : col[j] + off[q];
          e[q][j] = T.uc[i];
          tv[q][j] = T.t[i];
        }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (f0 + q < n) {
#pragma unroll
          for (int j = 0; j < NC; j++) {
            bad[j] |= (int)(e[q][j] >> 16) != code;
            if (MODE == MODE_REFINE) bad[j] |= (e[q][j] & 0xffffu) != 0u;
            else s[j] += tv[q][j];
          }
        }
      }
    }
  } else {
    for (int x = 0; x < w; x++) {
      int lyi[NC];
#pragma unroll
      for (int j = 0; j < NC; j++) lyi[j] = lyi0[j];
      for (int y = 0; y < h; y++) {
        uint32_t e[NC];
        float tv[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) {
          const int i = bad[j] ? 0 : T.at(lxi[j], lyi[j]);
          e[j] = T.uc[i];
          tv[j] = T.t[i];
        }
#pragma unroll
        for (int j = 0; j < NC; j++) {
          bad[j] |= (int)(e[j] >> 16) != code;
          if (MODE == MODE_REFINE) bad[j] |= (e[j] & 0xffffu) != 0u;
          else s[j] += tv[j];
          if (++lyi[j] == g.yh) lyi[j] = 0;
        }
      }
#pragma unroll
      for (int j = 0; j < NC; j++)
        if (++lxi[j] == g.xh) lxi[j] = 0;
    }
  }
#pragma unroll
  for (int j = 0; j < NC; j++) {
    float pc = (MODE == MODE_REFINE) ? 0.0f : s[j];
    cst[j] = bad[j] ? kInf : cst[j] + pc;
  }
  if (tech) {
    // cost = disp + penalty + pin_vio + edge_spacing (AdmmKernels.hpp:808-830),
    // the last two only where the penalty is finite
    const float pin_unit = a.tech_penalty * float(a.rhg);
    for (int j = 0; j < NC; j++) {
      if (bad[j]) continue;
      int nx, ny;
      new_pos(a, dir[j], w, h, cx, cy, nx, ny);
      const bool rot = will_rotate(h, T.vdd[lyi0[j]], gnd);
      const int npv = a.cmacro ? pin_vio_tile(a, g, T, tc.p0, tc.p1, w, h, rot, g.lx(nx), lyi0[j]) : 0;
      const float pin = pin_unit * float(npv);
      const float edge = edge_cost_tile(a, g, T, rot ? tc.rt : tc.lt, rot ? tc.lt : tc.rt, w, h, nx, ny);
      cst[j] = (cst[j] + pin) + edge;
    }
  }
}

// dual_update_functor_with_step_size (AdmmKernels.hpp:26-42): the literal 0.5
// promotes the quadratic term to double before the float store.
__device__ __forceinline__ float dual_step(int ui, float l, float rho, float step, float neg) {
  float m = step * dual_delta(ui, rho, neg);
  return fmax_ref(l + m, 0.0f);
}

// Dual segment column of x: the cuts lie where x + hint/3 * tf (tf = 0, 1, 2)
// is a multiple of hint, i.e. at x mod hint in {0, hint - 2 (hint/3),
// hint - hint/3} (the host's tables xseg / yseg hold the same numbers).
__device__ __forceinline__ int seg_axis(int x, int hint) {
  const int q = x / hint, r = x - q * hint, t = hint / 3;
  return 3 * q + (r >= hint - 2 * t ? 1 : 0) + (r >= hint - t ? 1 : 0);
}
// Dual segment of pixel (x, y).
__device__ __forceinline__ int seg_of(const Args& a, int x, int y) {
  return seg_axis(y, a.y_hint) * a.nxs + seg_axis(x, a.x_hint);
}

// The dual updates iterations from..to (inclusive) would have made to a
// pixel whose partition took no primal step in them: its usage u is then 0
// or 1 and unchanged; u = 1 adds zero, u = 0 lowers the price towards 0,
// where it stays.  The same dual_step, iteration by iteration, so the bits
// are those of the eager update.
// For usage 0 a step is l <- max(l + m0[k], 0) with m0[k] = step_k * delta(0)
// <= 0 (Args::m0_hist); the increments of iterations win0 .. win0 + kCatchWin
// - 1 may come from an LDS copy `win`.
constexpr int kCatchWin = 64;
__device__ __forceinline__ float dual_catch_up(const Args& a, int u, float l, int from, int to,
                                               const float* win = nullptr, int win0 = 0) {
  if (u != 0 || !(l > 0.0f)) return l;
  int k = from;
  if (win) {
    for (; k <= to && k < win0; k++) l = fmax_ref(l + a.m0_hist[k], 0.0f);
    const int end = min(to, win0 + kCatchWin - 1);
#pragma unroll 4
    for (; k <= end; k++) l = fmax_ref(l + win[k - win0], 0.0f);
  }
  for (; k <= to; k++) l = fmax_ref(l + a.m0_hist[k], 0.0f);
  return l;
}

// LDS copy of the increments m0[w0 .. w0 + kCatchWin) (entries before
// iteration 0 or from `last` on are not needed and left 0); every thread of
// the workgroup passes a barrier after it.
__device__ __forceinline__ void load_catch_window(const Args& a, float* win, int w0, int last, int tid, int nt) {
  for (int j = tid; j < kCatchWin; j += nt) {
    const int k = w0 + j;
    win[j] = (k >= 0 && k <= last) ? a.m0_hist[k] : 0.0f;
  }
  __syncthreads();
}

// Staging of the partition's pixel tile by the NW waves of the workgroup,
// threads walking the tile in order with incremental (la, lb) = (i % xh,
// i / xh): consecutive threads read consecutive pixels of a grid row;
// 16 (one wave) or 8 (four waves) pixels in flight per thread.  Returns
// true if a usage does not fit the 16-bit LDS packing.
// With catch_up (lazy duals, primal step of iteration `iter`), prices that
// missed dual updates are brought up to the end of iteration iter - 1.
template <int NW>
__device__ __forceinline__ bool stage_tile(const Args& a, const PartGeo& g, const Tile& T, float rho,
                                           bool standard, bool& over, int& cmin, int& cmax, int iter,
                                           bool catch_up, const float* win) {
  const float neg = -1.0f / rho;
  constexpr int NT = NW * kWave;
  const int tid = threadIdx.x;
  const int npix = g.xh * g.yh;
  const int step_b = NT / g.xh, step_a = NT - step_b * g.xh;
  int lb = tid / g.xh, la = tid - lb * g.xh;
  bool range_err = false;
  constexpr int U = NW > 1 ? 8 : 16;
  for (int i0 = 0; i0 < npix; i0 += U * NT) {
    int u[U], code[U], li[U], stamp[U];
    float dl[U];
    bool in[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      int xx = g.gx(la), yy = g.gy(lb);
      bool ok = (i0 + q * NT + tid < npix) && xx >= 0 && yy >= 0;
      // unconditional loads at a clamped index, then select
      const unsigned p = ok ? pidx(a, xx, yy) : 0u;
      li[q] = T.at(la, lb);
      in[q] = ok;
      const int uu = a.usage[p];
      const int cc = (int)(uint16_t)a.pcode[p];
      const float dd = a.dual[p];
      stamp[q] = catch_up && ok ? a.seg_stamp[seg_of(a, xx, yy)] : iter - 1;
      u[q] = ok ? uu : 0;
      code[q] = ok ? cc : 0;  // code 0: not placeable
      dl[q] = ok ? dd : 0.0f;
      la += step_a; lb += step_b;
      if (la >= g.xh) { la -= g.xh; lb++; }
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      int i = i0 + q * NT + tid;
      if (in[q] && code[q] > 0) {
        cmin = min(cmin, code[q]);
        cmax = max(cmax, code[q]);
      }
      if (i < npix) {
        if (stamp[q] < iter - 1)
          dl[q] = dual_catch_up(a, u[q], dl[q], stamp[q] + 1, iter - 1, win, iter - kCatchWin);
        range_err |= u[q] > 0xffff;
        over |= u[q] > 1;
        T.uc[li[q]] = ((uint32_t)code[q] << 16) | (uint32_t)(u[q] & 0xffff);
        // a pixel no cell may use (code 0) carries +inf: any footprint over
        // it sums to +inf, the cost of an invalid candidate
        T.t[li[q]] = code[q] == 0 ? kInf : pen_term(u[q], dl[q], rho, standard);
        T.lam[li[q]] = dl[q];
      }
    }
  }
  return range_err;
}

// key of computeKeyArrayForSortingKernel (SortingKernels.hpp:9-55) from the LDS usage
__device__ __forceinline__ int sort_key(const Args& a, const PartGeo& g, const Tile& T, int c, int4 rc) {
  int cr = 0, co = 10, ca = 1;
  if (a.sort_policy == 0) { cr = 1; co = 0; ca = 0; }
  int w = rec_w(rc.y), h = rec_h(rc.y);
  int ovf = co ? ovf_tile(g, T, rec_x(rc.x), rec_y(rc.x), w, h) : 0;
  return cr * (cr ? a.rand_prio[c] : 0) + co * ovf + ca * w * h;
}


// ---- the descending network on slots held lane-major -------------------
// Slot i = lane*RN + r lives in register r of lane `lane` (n = 64*RN slots, or
// n < 64 with RN = 1).  A step of distance j < RN pairs two registers of one
// lane; j >= RN pairs register r of lanes l and l ^ (j / RN).  The small
// distances, the most frequent ones in a bitonic network, thus need no
// cross-lane traffic.  PK: a key and its slot packed in one word
// ((key + 1) << 10 | slot, pads 0 << 10 | slot) compared on the key bits, so a
// cross-lane step moves one word; otherwise key and slot move separately.
// Compare-exchange semantics are those of AdmmKernels.hpp:899-945: the pair
// (lo, lo ^ j) is ordered descending if (lo & k) == 0, and swapped only if
// strictly out of order.
// v of lane (lane ^ d) for d = 1, 2, 4, 8 by DPP within a row of sixteen
// lanes (row_shl:N reads lane + N, row_shr:N lane - N, row_ror:8 lane ^ 8);
// other distances through the LDS crossbar
template <int D>
__device__ __forceinline__ int xor_lanes(int v) {
  if constexpr (D == 1) return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  else if constexpr (D == 2) return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]
  else if constexpr (D == 4) {
    const int up = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xF, 0xF, false);  // row_shl:4
    const int dn = __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    return (threadIdx.x & 4) ? dn : up;
  } else if constexpr (D == 8) return __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false);  // row_ror:8
  else return __shfl_xor(v, D);
}

// minimum and maximum over the wave, in every lane
__device__ __forceinline__ void wave_min_max(int& mn, int& mx) {
  mn = min(mn, xor_lanes<1>(mn)); mx = max(mx, xor_lanes<1>(mx));
  mn = min(mn, xor_lanes<2>(mn)); mx = max(mx, xor_lanes<2>(mx));
  mn = min(mn, xor_lanes<4>(mn)); mx = max(mx, xor_lanes<4>(mx));
  mn = min(mn, xor_lanes<8>(mn)); mx = max(mx, xor_lanes<8>(mx));
  mn = min(mn, xor_lanes<16>(mn)); mx = max(mx, xor_lanes<16>(mx));
  mn = min(mn, xor_lanes<32>(mn)); mx = max(mx, xor_lanes<32>(mx));
}

// comparison key of a network word: the key bits of a packed word, or the
// int with its sign flipped so that unsigned order is int order
template <bool PK>
__device__ __forceinline__ unsigned net_cmp(int v) {
  return PK ? ((unsigned)v >> 10) : ((unsigned)v ^ 0x80000000u);
}

// Pair (lo, lo + J) of one lane's registers.  Descending pairs ((lo & k) ==
// 0) swap if key(lo) < key(hi), ascending ones if key(hi) < key(lo): with
// x, y the two keys in the order the pair's direction gives, swap iff x < y.
template <int RN, int J, bool PK>
__device__ __forceinline__ void net_step_lane(int k, int (&key)[RN], int (&slot)[RN]) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int r = 0; r < RN; r++) {
    const int rp = r ^ J;
    if (rp > r) {
      const bool desc = ((lane * RN + r) & k) == 0;
      const int kl = key[r], kh = key[rp];
      const unsigned cl = net_cmp<PK>(kl), ch = net_cmp<PK>(kh);
      const bool sw = (desc ? cl : ch) < (desc ? ch : cl);
      key[r] = sw ? kh : kl;
      key[rp] = sw ? kl : kh;
      if (!PK) {
        const int sl = slot[r], sh = slot[rp];
        slot[r] = sw ? sh : sl;
        slot[rp] = sw ? sl : sh;
      }
    }
  }
}

// Register r of lanes l and l ^ D (slots i and i ^ j, j = D * RN).  Both
// slots of a pair have the same bit k (k > j), so the pair is descending iff
// (i & k) == 0; this lane holds the pair's low slot iff (i & j) == 0.  It
// takes the partner's word iff the pair swaps: for the low slot of a
// descending pair iff mine < other, for its high slot iff other < mine, and
// the reverse for an ascending pair -- x < y with (x, y) = (mine, other) when
// both tests agree, else (other, mine).
template <int RN, int D, bool PK>
__device__ __forceinline__ void net_step_cross(int k, int j, int (&key)[RN], int (&slot)[RN]) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int r = 0; r < RN; r++) {
    const int i = lane * RN + r;
    const bool t = ((i & k) == 0) == ((i & j) == 0);
    const int ok = xor_lanes<D>(key[r]);
    const unsigned cm = net_cmp<PK>(key[r]), co = net_cmp<PK>(ok);
    const bool take = (t ? cm : co) < (t ? co : cm);
    key[r] = take ? ok : key[r];
    if (!PK) {
      const int os = xor_lanes<D>(slot[r]);
      slot[r] = take ? os : slot[r];
    }
  }
}

template <int RN, bool PK>
__device__ __forceinline__ void net_sort(int n, int (&key)[RN], int (&slot)[RN]) {
  const int lane = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < RN) {
        if (j == 1) net_step_lane<RN, 1, PK>(k, key, slot);
        if (RN > 2 && j == 2) net_step_lane<RN, (RN > 2 ? 2 : 1), PK>(k, key, slot);
        if (RN > 4 && j == 4) net_step_lane<RN, (RN > 4 ? 4 : 1), PK>(k, key, slot);
        if (RN > 8 && j == 8) net_step_lane<RN, (RN > 8 ? 8 : 1), PK>(k, key, slot);
      } else {
        // register r of lanes l and l ^ (j / RN), j / RN <= 32
        switch (j / RN) {
          case 1: net_step_cross<RN, 1, PK>(k, j, key, slot); break;
          case 2: net_step_cross<RN, 2, PK>(k, j, key, slot); break;
          case 4: net_step_cross<RN, 4, PK>(k, j, key, slot); break;
          case 8: net_step_cross<RN, 8, PK>(k, j, key, slot); break;
          case 16: net_step_cross<RN, 16, PK>(k, j, key, slot); break;
          default: net_step_cross<RN, 32, PK>(k, j, key, slot); break;
        }
      }
    }
  }
}

// Processing order of one partition with the sort in registers (npad <=
// 64*RN; slots at or beyond npad are pads that only meet pads): keys of the
// cells in ascending id order (the grouping's order, that of
// thrust::stable_sort_by_key), the reference's descending bitonic network on
// npad slots, then the ids read back by slot.  Writes the order to
// grouped[off..] (ids) and rec_b[off..] (records).  The network moves packed
// words (key + 1) << 10 | slot; keys with key + 1 >= 2^22 (or < 0) are first
// replaced by their rank (the number of smaller keys, ties equal), which the
// network orders the same way -- in the instance of 16 registers only: the
// others then return false having written nothing.  The keys are computed in
// a loop that is not unrolled, so the kernel holds one copy of the overflow
// sums per instance.
template <int RN>
__device__ __forceinline__ bool order_partition_reg(const Args& a, const PartGeo& g, const Tile& T, int npad,
                                                    const int* src, int off, int count, const int4* cellrec
#ifdef LG_PHASE_TIMERS
                                                    , int* ph_
#endif
                                                    ) {
  const int lane = threadIdx.x;
#ifdef LG_PHASE_TIMERS
  unsigned long long t0_ = wall_clock64();
#endif
  int* ids = a.grouped + off;
  int4* recb = reinterpret_cast<int4*>(a.rec_b) + off;
  // the cells in ascending id order (the input order of the reference's
  // sort): the source ids through a descending network on -id (pads
  // INT_MIN sort last), stored to ids[]; loads at a clamped slot, selected
  // afterwards (slots >= count are pads)
  int val[RN];
  {
    int kk[RN], ss[RN];
#pragma unroll
    for (int r = 0; r < RN; r++) {
      const int i = lane * RN + r;
      const int v = src[min(i, count - 1)];
      kk[r] = i < count ? -v : INT_MIN;
      ss[r] = 0;
    }
    net_sort<RN, false>(npad, kk, ss);
    // pads take the last cell's id (slot count - 1)
    const int ls = count - 1, ll = ls / RN, lr = ls - ll * RN;
    int last = 0;
#pragma unroll
    for (int q = 0; q < RN; q++) {
      const int v = __shfl(-kk[q], ll);
      if (q == lr) last = v;
    }
    // (ids clamped into the cell range: the loads below stay in bounds)
#pragma unroll
    for (int r = 0; r < RN; r++) val[r] = min(max(lane * RN + r < count ? -kk[r] : last, 0), a.N - 1);
    if (RN > 8) {
      // the wide instance gathers by slot through memory
#pragma unroll
      for (int r = 0; r < RN; r++)
        if (lane * RN + r < count) ids[lane * RN + r] = val[r];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
  }
  // records loaded together, except in the wide instance, which loads each
  // when its key is computed (sixteen int4 would hold 64 registers)
  constexpr int RC = RN > 8 ? 1 : RN;
  int4 rc[RC];
#pragma unroll
  for (int r = 0; r < RC; r++) rc[r] = cellrec[val[r]];
  int key[RN];
  int kmin = INT_MAX, kmax = INT_MIN;
#pragma unroll 1
  for (int r = 0; r < RN; r++) {
    int v = val[0];
    int4 rr = rc[0];
#pragma unroll
    for (int q = 1; q < RN; q++)
      if (q == r) v = val[q];
    if (RN > 8) rr = cellrec[v];
    else {
#pragma unroll
      for (int q = 1; q < RC; q++)
        if (q == r) rr = rc[q];
    }
    const int i = lane * RN + r;
    const int kk = i < count ? sort_key(a, g, T, v, rr) : INT_MIN;
    if (i < count) {
      kmin = min(kmin, kk);
      kmax = max(kmax, kk);
    }
#pragma unroll
    for (int q = 0; q < RN; q++)
      if (q == r) key[q] = kk;
  }
  wave_min_max(kmin, kmax);
  if (kmin < 0 || kmax >= a.pack_limit) {
    // the instance of 16 registers handles such keys
    if (RN < 16) return false;
    // rank of each cell key among the cell keys (pads stay INT_MIN)
    int rk[RN];
#pragma unroll
    for (int r = 0; r < RN; r++) rk[r] = 0;
#pragma unroll 1
    for (int j = 0; j < count; j++) {
      const int src = j / RN, q = j - src * RN;
      int kq = key[0];
#pragma unroll
      for (int u = 1; u < RN; u++)
        if (u == q) kq = key[u];
      const int ks = __shfl(kq, src);
#pragma unroll
      for (int r = 0; r < RN; r++) rk[r] += ks < key[r] ? 1 : 0;
    }
#pragma unroll
    for (int r = 0; r < RN; r++)
      if (lane * RN + r < count) key[r] = rk[r];
  }
#ifdef LG_PHASE_TIMERS
  unsigned long long t1_ = wall_clock64();
  ph_[5] = (int)(t1_ - t0_);
#endif
  int slot[RN];
#pragma unroll
  for (int r = 0; r < RN; r++) {
    slot[r] = lane * RN + r;
    key[r] = (int)(((unsigned)(key[r] == INT_MIN ? 0 : key[r] + 1) << 10) | (unsigned)slot[r]);
  }
  net_sort<RN, true>(npad, key, slot);
#pragma unroll
  for (int r = 0; r < RN; r++) slot[r] = key[r] & 1023;
#ifdef LG_PHASE_TIMERS
  unsigned long long t2_ = wall_clock64();
  ph_[6] = (int)(t2_ - t1_);
#endif
  // pads sort behind every cell, so slots < count map to cells; the id of
  // slot s lives in register s % RN of lane s / RN
  if (RN <= 8) {
    int nv[RN];
#pragma unroll
    for (int r = 0; r < RN; r++) {
      const int sl = min(slot[r], count - 1), sr = sl - (sl / RN) * RN;
      nv[r] = 0;
#pragma unroll
      for (int q = 0; q < RN; q++) {
        const int v = __shfl(val[q], sl / RN);
        if (q == sr) nv[r] = v;
      }
    }
#pragma unroll
    for (int r = 0; r < RN; r++) val[r] = nv[r];
  } else {
#pragma unroll
    for (int r = 0; r < RN; r++) val[r] = ids[min(slot[r], count - 1)];
  }
  if (RN > 8) {
#pragma unroll
    for (int r = 0; r < RN; r++) {
      const int i = lane * RN + r;
      if (i < count) {
        ids[i] = val[r];
        recb[i] = cellrec[val[r]];
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < RC; r++) rc[r] = cellrec[val[r]];
#pragma unroll
    for (int r = 0; r < RC; r++) {
      const int i = lane * RN + r;
      if (i < count) {
        ids[i] = val[r];
        recb[i] = rc[r];
      }
    }
  }
#ifdef LG_PHASE_TIMERS
  ph_[7] = (int)(wall_clock64() - t2_);
#endif
  return true;
}

// The network on n <= 256 slots held one per thread of a four-wave
// workgroup (slot = thread): distances below 64 within a wave (DPP / LDS
// crossbar), 64 and 128 through the LDS words xw[256].  Same pairs,
// directions and swap rule as net_step_cross.
template <bool PK>
__device__ __forceinline__ void coop_net(int n, int& w, int* xw) {
  const int tid = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      int o;
      if (j >= kWave) {
        xw[tid] = w;
        __syncthreads();
        o = xw[tid ^ j];
        __syncthreads();
      } else {
        switch (j) {
          case 1: o = xor_lanes<1>(w); break;
          case 2: o = xor_lanes<2>(w); break;
          case 4: o = xor_lanes<4>(w); break;
          case 8: o = xor_lanes<8>(w); break;
          case 16: o = xor_lanes<16>(w); break;
          default: o = xor_lanes<32>(w); break;
        }
      }
      const bool t = ((tid & k) == 0) == ((tid & j) == 0);
      const unsigned cm = net_cmp<PK>(w), co = net_cmp<PK>(o);
      w = ((t ? cm : co) < (t ? co : cm)) ? o : w;
    }
  }
}

// Processing order of a partition of npad <= 256 slots by the four waves,
// one slot per thread: the same two networks and keys as
// order_partition_reg.  sids[256] (LDS) holds the ids in ascending order,
// xw[256] the network's exchange words, xch[32..48) the key range of the
// waves.  Returns false, having written nothing, if the keys do not pack.
__device__ __forceinline__ bool order_partition_coop(const Args& a, const PartGeo& g, const Tile& T, int npad,
                                                     const int* src, int off, int count, const int4* cellrec,
                                                     int* sids, int* xw, int* xch
#ifdef LG_PHASE_TIMERS
                                                     , int* ph_
#endif
                                                     ) {
  constexpr int NW = 4;
#ifdef LG_PHASE_TIMERS
  unsigned long long t0_ = wall_clock64();
#endif
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  int* ids = a.grouped + off;
  int4* recb = reinterpret_cast<int4*>(a.rec_b) + off;
  // the cells in ascending id order: ids are distinct, so an id's slot is
  // its rank, the number of smaller ids (every thread counts over the
  // bucket in LDS, broadcast reads; INT_MAX past the count)
  const int v = tid < count ? src[tid] : INT_MAX;
  xw[tid] = v;
  __syncthreads();
  if (tid < count) {
    int rank = 0;
    for (int q = 0; q < count; q += 4) {
      const int4 u = *reinterpret_cast<const int4*>(xw + q);
      rank += (u.x < v ? 1 : 0) + (u.y < v ? 1 : 0) + (u.z < v ? 1 : 0) + (u.w < v ? 1 : 0);
    }
    sids[rank] = v;
  }
  __syncthreads();
  // ids clamped into the cell range
  const int id = tid < count ? min(max(sids[tid], 0), a.N - 1) : 0;
#ifdef LG_PHASE_TIMERS
  unsigned long long t1_ = wall_clock64();
  ph_[5] = (int)(t1_ - t0_);
#endif
  int key = INT_MIN, kmin = INT_MAX, kmax = INT_MIN;
  if (tid < count) {
    key = sort_key(a, g, T, id, cellrec[id]);
    kmin = kmax = key;
  }
  wave_min_max(kmin, kmax);
  if (lane == 0) {
    xch[32 + wave] = kmin;
    xch[40 + wave] = kmax;
  }
  __syncthreads();
#pragma unroll
  for (int w2 = 0; w2 < NW; w2++) {
    kmin = min(kmin, xch[32 + w2]);
    kmax = max(kmax, xch[40 + w2]);
  }
  if (kmin < 0 || kmax >= a.pack_limit) return false;
#ifdef LG_PHASE_TIMERS
  unsigned long long t2_ = wall_clock64();
  ph_[6] = (int)(t2_ - t1_);
#endif
  int word = (int)(((unsigned)(key == INT_MIN ? 0 : key + 1) << 10) | (unsigned)tid);
  coop_net<true>(npad, word, xw);
#ifdef LG_PHASE_TIMERS
  ph_[7] = (int)(wall_clock64() - t2_);
#endif
  // pads sort behind every cell: slots below count hold cells
  if (tid < count) {
    const int c = sids[word & 1023];
    ids[tid] = c;
    recb[tid] = cellrec[c];
  }
  return true;
}

// The same with the sorts in LDS (npad > 1024, or the merge / radix options).
__device__ void order_partition_lds(const Args& a, const PartGeo& g, const Tile& T, int npad, int npad_hi,
                                    const int* src, int off, int count, const int4* cellrec, int* keys, int* slots,
                                    int* aux) {
  const int lane = threadIdx.x;
  int* ids = a.grouped + off;
  int4* recb = reinterpret_cast<int4*>(a.rec_b) + off;
  // the cells in ascending id order (the input order of the reference's
  // sort): a descending network on -id, pads INT_MIN last
  for (int i = lane; i < npad; i += kWave) {
    keys[i] = i < count ? -src[i] : INT_MIN;
    slots[i] = 0;
  }
  __syncthreads();
  lds_bitonic_desc(npad, keys, slots);
  for (int i = lane; i < npad; i += kWave) {
    if (i < count) {
      int c = -keys[i];
      keys[i] = sort_key(a, g, T, c, cellrec[c]);
      slots[i] = c;
    } else {
      keys[i] = INT_MIN;
      slots[i] = -1;
    }
  }
  __syncthreads();
  if (a.sort_algo == 0 || a.sort_algo == 1) lds_merge_desc(npad, keys, slots, aux, aux + npad_hi);
  else if (a.sort_algo == 4) lds_stable_desc(npad, count, keys, slots, aux);
  else lds_bitonic_desc(npad, keys, slots);
  for (int i = lane; i < count; i += kWave) {
    int c = slots[i];
    ids[i] = c;
    recb[i] = cellrec[c];
  }
}

// Appends partition k to the large class's list of this iteration; that
// launch runs after the current one and reads the count then.
__device__ __forceinline__ void defer_partition(const Args& a, int k) {
  const int pos = atomicAdd(&a.ctl->n_work[1], 1);
  a.work[a.K + pos] = k;
}

// Or of a predicate over the workgroup's NW waves; four-wave workgroups
// exchange through LDS slots x[0..NW), each call site with slots of its own.
template <int NW>
__device__ __forceinline__ bool block_any(bool v, int* x, int lane, int wave) {
  const bool any = __ballot(v) != 0ull;
  if (NW == 1) return any;
  if (lane == 0) x[wave] = any ? 1 : 0;
  __syncthreads();
  int r = 0;
#pragma unroll
  for (int w2 = 0; w2 < NW; w2++) r |= x[w2];
  return r != 0;
}

// activity flag, under the next iteration's fold offset, of the partition
// holding pixel (x, y); set for every pixel with usage > 1 that a primal step
// writes (only partitions that take the step can hold such pixels)
// (computed like the colmap / rowmap tables, without their loads)
__device__ __forceinline__ void flag_next(const Args& a, int iter, int x, int y) {
  const int tn = triple_fold(a, iter + 1);
  const int col = ((x + a.x_hint / 3 * tn) / a.x_hint) % a.ncol;
  const int row = ((y + a.y_hint / 3 * tn) / a.y_hint) % a.nrow;
  a.act_flags[(size_t)((iter + 1) & 1) * a.K + col * a.nrow + row] = 1;
}

// Lazy duals: the partition's prices hold iteration `iter`'s update now.
// Writes the stamp of every dual segment of the tile once (at the segment's
// first column and row inside the tile); called after a barrier that
// follows every stamp read of the workgroup.
__device__ __forceinline__ void stamp_partition(const Args& a, const PartGeo& g, int iter, int tid, int nt) {
  const int npix = g.xh * g.yh;
  for (int i = tid; i < npix; i += nt) {
    const int lb = i / g.xh, la = i - lb * g.xh;
    const int x = g.gx(la), y = g.gy(lb);
    if (x < 0 || y < 0) continue;
    // (x = 0 / y = 0 start a segment; through the wrap-around of the
    // colouring they can sit inside the tile, after columns outside the grid)
    const bool fx = la == 0 || x == 0 || g.gx(la - 1) != x - 1 || seg_axis(x - 1, a.x_hint) != seg_axis(x, a.x_hint);
    const bool fy = lb == 0 || y == 0 || g.gy(lb - 1) != y - 1 || seg_axis(y - 1, a.y_hint) != seg_axis(y, a.y_hint);
    if (fx && fy) a.seg_stamp[seg_of(a, x, y)] = iter;
  }
}

// doPrimalUpdateKernel (AdmmKernels.hpp:637-897) / doFinalRefineKernel
// (AdmmKernels.hpp:1176-1416) for one partition, plus -- in primal mode -- the
// dual update and the overflow count of the partition's pixels.
//
// One 64-lane wavefront owns one partition.  If none of its cells needs a move
// (no overflow in its footprint, correct rail) the primal step is a no-op and
// only the dual update streams over its pixels.  Otherwise the pixels are
// staged in LDS, the processing order is sorted, and the cells are visited in
// that order; whether a cell is skipped is evaluated for a window of 64
// upcoming cells at once and re-evaluated after every move that touches its
// footprint, which is exactly the sequential decision because a skipped cell
// changes nothing.
//
// npad_lo..npad_hi selects the launch (LDS size class); LDS_SORT launches keep
// the sort buffers in LDS, the others sort in registers (npad <= 1024).
template <int MODE, int LDS_SORT, int NW>
__device__ __forceinline__ void partition_body(const Args& a, int bx, int npad_lo, int npad_hi) {
  static_assert(NW == 1 || (NW == 4 && LDS_SORT == 0 && MODE == MODE_PRIMAL), "four waves: primal, register sorts");
  constexpr int NT = NW * kWave;
  // Four-wave launches sort up to 8 * 64 slots in registers.  A larger
  // partition, or one whose keys are too wide to pack, is appended to the
  // list of the large class, whose one-wave launch (LDS sorts) comes next.
  constexpr bool kDefer = NW > 1;
  Ctl* ctl = a.ctl;
  if (MODE == MODE_PRIMAL && ctl->done) return;
  // the primal step takes its partitions from the work list of its class;
  // the final refine visits every partition
  const int cls = npad_lo > 0 ? 1 : 0;
  if (MODE == MODE_PRIMAL && (int)bx >= ctl->n_work[cls]) return;
  if (MODE != MODE_PRIMAL && ctl->npad < npad_lo) return;  // no partition of this size class
  const int k = MODE == MODE_PRIMAL ? a.work[cls * a.K + bx] : (int)bx;
  const bool adaptive_now = (MODE == MODE_PRIMAL) && a.use_adaptive && ctl->exceptional;
  const int iter = MODE == MODE_PRIMAL ? ctl->iter : ctl->conv_iter;
  const int off = a.off[k];
  const int count = adaptive_now ? 0 : a.pcnt[k];
  // Sort size of this partition.  The reference sorts every partition on
  // npad = pow2(largest partition) slots, padding with INT_MIN; a bitonic
  // network whose real keys fill the first c slots sorts them exactly as the
  // network on pow2(c) slots does (the later merge stages only compare
  // ordered runs with padding, which never swaps), and the merge / radix
  // options are stable, so each partition sorts pow2(count) slots.
  int npad = 1;
  while (npad < count) npad <<= 1;
  if (npad < npad_lo || npad > npad_hi) return;
  if (kDefer && npad > 8 * kWave) {
    if (threadIdx.x == 0) defer_partition(a, k);
    return;
  }
  PHASE_START();
  SUB_DECL();
  const int tid = threadIdx.x;
  const int lane = NW == 1 ? tid : (tid & (kWave - 1));
  const int wave = NW == 1 ? 0 : (tid >> 6);
  PartGeo g;
  const int tf = triple_fold(a, iter);
  g.xh = a.x_hint; g.yh = a.y_hint; g.ncol = a.ncol; g.nrow = a.nrow;
  g.oxs = g.xh / 3 * tf; g.oys = g.yh / 3 * tf;
  g.col = k / a.nrow; g.row = k - g.col * a.nrow;
  g.init(a.x_size, a.y_size);
  int4* cellrec = reinterpret_cast<int4*>(a.cellrec);
  const int npix = g.xh * g.yh;
  const int step_b = NT / g.xh, step_a = NT - step_b * g.xh;

  const float rho = ctl->rho;
  const float step = ctl->step;
  const float neg = -1.0f / rho;
  const bool pert = (MODE == MODE_PRIMAL) && iter < a.iter_perturbation_on;
  const bool standard = (MODE == MODE_PRIMAL) && a.standard_admm;
  int* ids = a.grouped + off;  // the processing order, once sorted
  // the partition's cells: its bucket (arbitrary order), or the order the
  // last primal step left for the final refine
  const bool from_bucket = MODE == MODE_PRIMAL || a.sorted[k] == 0;
  const int* src = from_bucket ? a.bucket + (size_t)k * kMaxNpad : ids;

  extern __shared__ __align__(16) unsigned char lds_raw[];
  Tile T;
  T.rs = g.xh <= kRowStride ? kRowStride : g.xh;
  const int tsz = (T.rs * g.yh + 3) & ~3;
  T.uc = reinterpret_cast<uint32_t*>(lds_raw);                 // [rs * yh]
  T.t = reinterpret_cast<float*>(T.uc + tsz);                  // [rs * yh]
  T.lam = T.t + tsz;                                           // [rs * yh]
  T.vdd = reinterpret_cast<int*>(T.lam + tsz);                 // [yh]
  int* after_tile = T.vdd + ((g.yh + 3) & ~3);
  T.eL = T.eR = nullptr;
  T.pm = nullptr;
  if (NW == 1 && a.tech) {  // edge maps and power metal of the tile
    T.eL = reinterpret_cast<uint32_t*>(after_tile);            // [rs * yh]
    T.eR = T.eL + tsz;                                         // [rs * yh]
    T.pm = reinterpret_cast<int8_t*>(T.eR + tsz);              // [rs * yh] bytes
    after_tile = reinterpret_cast<int*>(T.pm + tsz);           // tsz is a multiple of 4
  }
  // LDS sorts: key / slot / buffer arrays of lds_hi slots in LDS, or, for a
  // partition above kMaxNpadLds slots, in its own global-memory buffer
  const int lds_hi = npad_hi < kMaxNpadLds ? npad_hi : kMaxNpadLds;
  int* keys = after_tile;                                      // LDS sorts only
  int* slots = keys + lds_hi;
  int* aux = slots + lds_hi;
  int sort_hi = lds_hi;
  if (LDS_SORT && npad > kMaxNpadLds) {
    int hs = 0;
    if (tid == 0) hs = atomicAdd(&ctl->n_huge, 1);
    hs = __shfl(hs, 0);
    // every such partition holds more than kMaxNpadLds distinct cells, so
    // at most N / (kMaxNpadLds + 1) of them run at once: huge_slots
    if (hs >= a.huge_slots) {
      if (tid == 0) ctl->error = kErrPartitionTooDense;
      return;
    }
    keys = a.huge + (size_t)hs * 4 * kMaxNpad;
    slots = keys + kMaxNpad;
    aux = slots + kMaxNpad;
    sort_hi = kMaxNpad;
  }
  // exchange slots of the waves (four-wave launches, which have no LDS
  // sorts): block-wide or, code range, and the wave minima of the moves
  int* xch = keys;                                             // [64]
  unsigned* xmin = reinterpret_cast<unsigned*>(xch + 16);      // [2][NW]; xch[24]: deferred
  int* sids = xch + 64;                                        // [256] four-wave sorts
  int* xw = sids + 256;                                        // [256]
  (void)aux;

  // -- activity test: does any cell of this partition need a move?  With
  // every cell on a correct-parity row, a partition whose pixels all have
  // usage <= 1 has none (its cells' footprints lie inside it); that test is
  // fused into the tile staging below.  Otherwise each cell is checked.
  const bool fast_ok = MODE == MODE_PRIMAL && !pert;
  const bool cell_test = fast_ok && ctl->wrong_row > 0;
  if (fast_ok && !cell_test && !adaptive_now && a.act_flags[(size_t)(iter & 1) * a.K + k] == 0) {
    // no pixel above usage one: no cell moves, and k_dual_pass updates
    // this partition's prices
    if (tid == 0) a.sorted[k] = 0;
    return;
  }
  // lazy duals: the increments of the last kCatchWin iterations in LDS for
  // the prices this step brings up to date
  const bool lazy = MODE == MODE_PRIMAL && a.lazy;
  __shared__ float s_catch[kCatchWin];
  if (lazy) load_catch_window(a, s_catch, iter - kCatchWin, iter - 1, tid, NT);
  if (cell_test) {
    bool need = false;
    for (int i = tid; i < count; i += NT) {
      int4 rc = cellrec[src[i]];
      int cx = rec_x(rc.x), cy = rec_y(rc.x), h = rec_h(rc.y);
      need |= !correct_row(h, rec_gnd(rc.y), a.vdd_up[cy]) || ovf_hbm(a, cx, cy, rec_w(rc.y), h) != 0;
    }
    if (!block_any<NW>(need, xch, lane, wave)) {
      // primal step is a no-op: stream the dual update over the pixels
      unsigned long long ovf = 0;
      int lb = tid / g.xh, la = tid - lb * g.xh;
      for (int i = tid; i < npix; i += NT) {
        int xx = g.gx(la), yy = g.gy(lb);
        la += step_a; lb += step_b;
        if (la >= g.xh) { la -= g.xh; lb++; }
        if (xx < 0 || yy < 0) continue;
        const unsigned p = pidx(a, xx, yy);
        int ui = a.usage[p];
        float l = a.dual[p];
        if (a.lazy) {
          const int sg = seg_of(a, xx, yy);
          const int s = a.seg_stamp[sg];
          if (s < iter - 1) l = dual_catch_up(a, ui, l, s + 1, iter - 1, s_catch, iter - kCatchWin);
        }
        a.dual[p] = dual_step(ui, l, rho, step, neg);
        ovf += (unsigned long long)max(ui - 1, 0);
        if (ui > 1) flag_next(a, iter, xx, yy);
      }
      for (int o = 32; o > 0; o >>= 1) ovf += __shfl_down(ovf, o);
      if (lane == 0 && ovf) atomicAdd(&ctl->ovf_acc, ovf);
      if (tid == 0) a.sorted[k] = 0;
      if (a.lazy) {
        __syncthreads();
        stamp_partition(a, g, iter, tid, NT);
      }
      PHASE_MARK(4);
      PHASE_END();
      return;
    }
  }

  bool over = false;
  int cmin = 0x7fffffff, cmax = -1;
  if (stage_tile<NW>(a, g, T, rho, standard, over, cmin, cmax, iter, lazy, s_catch)) ctl->error = kErrUsageRange;
  // one placeable code on the tile's placeable pixels (the others carry
  // +inf penalty terms): candidate evaluation then needs no per-pixel code
  // test
  wave_min_max(cmin, cmax);
  if (NW > 1 && lane == 0) {
    xch[8 + wave] = cmin;
    xch[12 + wave] = cmax;
  }
  if (fast_ok && !cell_test && !block_any<NW>(over, xch + 4, lane, wave)) {
    // no cell can move: dual update from the staged tile (usage unchanged,
    // overflow zero)
    int lb = tid / g.xh, la = tid - lb * g.xh;
    for (int i = tid; i < npix; i += NT) {
      int xx = g.gx(la), yy = g.gy(lb);
      const int li = T.at(la, lb);
      la += step_a; lb += step_b;
      if (la >= g.xh) { la -= g.xh; lb++; }
      if (xx < 0 || yy < 0) continue;
      a.dual[pidx(a, xx, yy)] = dual_step((int)(T.uc[li] & 0xffffu), T.lam[li], rho, step, neg);
    }
    if (tid == 0) a.sorted[k] = 0;
    if (lazy) {
      if (NW > 1) __syncthreads();
      stamp_partition(a, g, iter, tid, NT);
    }
    PHASE_MARK(4);
    PHASE_END();
    return;
  }
  for (int b = tid; b < g.yh; b += NT) {
    int yy = g.gy(b);
    T.vdd[b] = yy >= 0 ? a.vdd_up[yy] : 0;
  }
  if (NW == 1 && a.tech) {
    // the tile's edge maps (no other partition writes them this step) and
    // power metal; pixels outside the grid carry no edges and no metal
    for (int i = tid; i < npix; i += NT) {
      const int lb = i / g.xh, la = i - lb * g.xh;
      const int xx = g.gx(la), yy = g.gy(lb);
      const int li = T.at(la, lb);
      const bool in = xx >= 0 && yy >= 0;
      const unsigned p = in ? pidx(a, xx, yy) : 0u;
      T.eL[li] = in ? a.edgeL[p] : kEdgeBias;
      T.eR[li] = in ? a.edgeR[p] : kEdgeBias;
      T.pm[li] = in ? a.pmetal[p] : (int8_t)-1;
    }
  }
  __syncthreads();
  if (NW > 1) {
#pragma unroll
    for (int w2 = 0; w2 < NW; w2++) {
      cmin = min(cmin, xch[8 + w2]);
      cmax = max(cmax, xch[12 + w2]);
    }
  }
  const int ucode = (MODE == MODE_PRIMAL && cmin == cmax && cmin > 0) ? cmin : -1;
  // rail of every tile row as one mask (rows < 64), for the candidate test
  const unsigned long long vmask = __ballot(lane < g.yh && T.vdd[lane < g.yh ? lane : 0] != 0);
  PHASE_MARK(0);
  PHASE_COUNT(8);

  if (count > 0) {
    int4* recs = reinterpret_cast<int4*>(a.rec_b) + off;
    if (from_bucket) {
      // an unsorted partition at the final refine moved nothing in the last
      // iteration, so its usage still is the one the order is defined on
      if (LDS_SORT) order_partition_lds(a, g, T, npad, sort_hi, src, off, count, cellrec, keys, slots, aux);
      else if (NW > 1 && npad <= NT) {
        // one slot per thread
#ifdef LG_PHASE_TIMERS
        const bool ok = order_partition_coop(a, g, T, npad, src, off, count, cellrec, sids, xw, xch, ph_);
#else
        const bool ok = order_partition_coop(a, g, T, npad, src, off, count, cellrec, sids, xw, xch);
#endif
        if (tid == 0) xch[24] = ok ? 0 : 1;
      } else if (wave == 0) {  // the register sorts run on the first wave
#ifdef LG_PHASE_TIMERS
#define ORDER_REG(RN) order_partition_reg<RN>(a, g, T, npad, src, off, count, cellrec, ph_)
#else
#define ORDER_REG(RN) order_partition_reg<RN>(a, g, T, npad, src, off, count, cellrec)
#endif
        // npad / 64 registers per lane, at least four; keys too wide to
        // pack go to the instance of 16 registers
        bool ok = false;
        if (npad <= 4 * kWave) ok = ORDER_REG(4);
        else if (npad == 8 * kWave) ok = ORDER_REG(8);
        if constexpr (kDefer) {
          if (tid == 0) xch[24] = ok ? 0 : 1;
        } else {
          if (!ok) ORDER_REG(16);
        }
#undef ORDER_REG
      }
      if (MODE == MODE_PRIMAL && tid == 0) a.sorted[k] = 1;
    } else {
      for (int i = tid; i < count; i += NT) recs[i] = cellrec[ids[i]];
    }
    __syncthreads();
    if (kDefer && xch[24] != 0) {
      if (tid == 0) defer_partition(a, k);
      return;
    }
    PHASE_MARK(1);

    const float thr = disp_threshold(a);
    // this lane's candidate moves d = d0 + j*64, j < NC, for the whole
    // kernel: four per lane on one wave, one per lane on four
    constexpr int NC = 4 / NW;
    const int d0 = wave * kWave + lane;
    int2 dir[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) {
      int d = d0 + j * kWave;
      dir[j] = d < kNumDirDefault ? c_dir_default[d] : make_int2(0, 0);
    }
    // -- the cells in order, 64 at a time ---------------------------------
    if constexpr (NW > 1) {
      // Wave 0 keeps the window, runs the skip tests, rips up and commits;
      // every wave evaluates its share of the candidates.  Per move wave 0
      // posts {next cell, its record} (or -1: window done) in one of two
      // LDS slots by pass parity, then two barriers: B1 (the post and the
      // rip-up are visible), B2 (the waves' candidate minima are posted).
      // Tile writes happen only on wave 0 between B2 and the next B1, while
      // the other waves read nothing of the tile.
      int pass = 0;
      for (int base = 0; base < count; base += kWave) {
        const int t = base + lane;
        int4 r = make_int4(0, 0, 0, 0);
        int id = -1;
        bool need = false;
        int start = 0;
        if (wave == 0) {
          if (t < count) { r = recs[t]; id = ids[t]; }
        }
        const int my_x = rec_x(r.x), my_y = rec_y(r.x), my_w = rec_w(r.y), my_h = rec_h(r.y);
        // skip test of computeCellOverflow + isPlacedInCorrectRow (AdmmKernels.hpp:722-738)
        auto skip_test = [&]() -> bool {
          if (t >= count) return false;
          if (pert) return true;
          bool ok_row = correct_row(my_h, rec_gnd(r.y), T.vdd[g.ly(my_y)]);
          return !ok_row || ovf_tile(g, T, my_x, my_y, my_w, my_h) != 0;
        };
        if (wave == 0) need = skip_test();
        int pid = 0;
        for (;;) {
          int* post = xch + 48 + 8 * (pass & 1);
          SUB_START();
          if (wave == 0) {
            const unsigned long long m = __ballot(need && lane >= start);
            if (m != 0ull) {
              const int first = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
              const int px = __builtin_amdgcn_readlane(r.x, first), py = __builtin_amdgcn_readlane(r.y, first);
              pid = __builtin_amdgcn_readlane(id, first);
              if (lane == 0) {
                post[0] = first;
                post[1] = px;
                post[2] = py;
                post[3] = __builtin_amdgcn_readlane(r.z, first);
                post[4] = __builtin_amdgcn_readlane(r.w, first);
              }
#ifdef LG_PHASE_TIMERS
              sub_acc_[5]++;
#endif
              footprint_add<1>(g, T, rec_x(px), rec_y(px), rec_w(py), rec_h(py), -1, rho, standard);  // rip-up
              SUB_MARK(11);
            } else if (lane == 0) {
              post[0] = -1;
            }
          }
          __syncthreads();  // B1
          pass++;
          const int first = post[0];
          if (first < 0) break;
          const int px = post[1], py = post[2], pox = post[3], poy = post[4];
          const int cx = rec_x(px), cy = rec_y(px);
          const int w = rec_w(py), h = rec_h(py), gnd = rec_gnd(py), code = rec_code(py);
          float disp_coeff = 1.0f;
          if (a.flag_iccad17) disp_coeff = float(a.N) / float(a.k_height_num[h - 1]);
          float cst[NC];
          const TechCell no_tech{0, 0, 0, 0};
          eval_candidates<MODE, NC>(a, g, T, disp_coeff, thr, cx, cy, w, h, code, gnd, pox, poy, d0, dir, cst,
                                    ucode, vmask, no_tech);
          SUB_MARK(12);
          // the tree-min as below: a DPP butterfly per wave, then the four
          // wave minima
          unsigned long long best = ~0ull;
#pragma unroll
          for (int j = 0; j < NC; j++) {
            unsigned slot = (unsigned)(d0 + j * kWave);
            unsigned rev = __brev(slot) >> 24;
            unsigned long long key = ((unsigned long long)__float_as_uint(cst[j]) << 32) | rev;
            best = key < best ? key : best;
          }
          unsigned bhi = (unsigned)(best >> 32), blo = (unsigned)best;
          key_min_step<0xB1, 0xf>(bhi, blo);
          key_min_step<0x4E, 0xf>(bhi, blo);
          key_min_step<0x141, 0xf>(bhi, blo);
          key_min_step<0x140, 0xf>(bhi, blo);
          key_min_step<0x142, 0xa>(bhi, blo);
          key_min_step<0x143, 0xc>(bhi, blo);
          if (lane == kWave - 1) {
            xmin[wave] = bhi;
            xmin[NW + wave] = blo;
          }
          __syncthreads();  // B2
          if (wave == 0) {
            unsigned bc_bits = xmin[0], bc_lo = xmin[NW];
#pragma unroll
            for (int w2 = 1; w2 < NW; w2++) {
              const unsigned ohi = xmin[w2], olo = xmin[NW + w2];
              const bool take = ohi < bc_bits || (ohi == bc_bits && olo < bc_lo);
              bc_bits = take ? ohi : bc_bits;
              bc_lo = take ? olo : bc_lo;
            }
            const int bslot = (int)(__brev(bc_lo & 0xffu) >> 24);
            int nx = cx, ny = cy;
            if (__uint_as_float(bc_bits) != kInf) {
              int2 dd = c_dir_default[bslot];
              new_pos(a, dd, w, h, cx, cy, nx, ny);
            }
            SUB_MARK(13);
            footprint_add<1>(g, T, nx, ny, w, h, +1, rho, standard);  // commit
            if (lane == 0) cellrec[pid].x = rec_pos(nx, ny);
            SUB_MARK(14);
            // a move changes usage only inside the old and new footprints
            start = first + 1;
            if (lane >= start && t < count && !pert) {
              bool hit_old = my_x < cx + w && cx < my_x + my_w && my_y < cy + h && cy < my_y + my_h;
              bool hit_new = my_x < nx + w && nx < my_x + my_w && my_y < ny + h && ny < my_y + my_h;
              if (hit_old || hit_new) need = skip_test();
            }
            SUB_MARK(15);
          }
        }
      }
    } else {
      for (int base = 0; base < count; base += kWave) {
        const int t = base + lane;
        int4 r = make_int4(0, 0, 0, 0);
        int id = -1;
        if (t < count) { r = recs[t]; id = ids[t]; }
        const int my_x = rec_x(r.x), my_y = rec_y(r.x), my_w = rec_w(r.y), my_h = rec_h(r.y);
        // skip test of computeCellOverflow + isPlacedInCorrectRow (AdmmKernels.hpp:722-738)
        auto skip_test = [&]() -> bool {
          if (t >= count) return false;
          if (MODE == MODE_PRIMAL) {
            if (pert) return true;
            bool ok_row = correct_row(my_h, rec_gnd(r.y), T.vdd[g.ly(my_y)]);
            return !ok_row || ovf_tile(g, T, my_x, my_y, my_w, my_h) != 0;
          }
          return my_h > 1;
        };
        bool need = skip_test();
        int start = 0;
        for (;;) {
          const unsigned long long m = __ballot(need && lane >= start);
          if (m == 0ull) break;
          SUB_START();
          const int first = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
          const int px = __builtin_amdgcn_readlane(r.x, first), py = __builtin_amdgcn_readlane(r.y, first);
          const int pox = __builtin_amdgcn_readlane(r.z, first), poy = __builtin_amdgcn_readlane(r.w, first);
          const int pid = __builtin_amdgcn_readlane(id, first);
          const int cx = rec_x(px), cy = rec_y(px);
          const int w = rec_w(py), h = rec_h(py), gnd = rec_gnd(py), code = rec_code(py);
          float disp_coeff = 1.0f;
          if (a.flag_iccad17) disp_coeff = float(a.N) / float(a.k_height_num[h - 1]);
  #ifdef LG_PHASE_TIMERS
          sub_acc_[5]++;
  #endif
          // the moved cell's technology: edge types and its master's pins
          TechCell tc{0, 0, 0, 0};
          const bool tech = NW == 1 && a.tech != 0;
          if (tech) {
            const int ce = a.cedge[pid];
            tc.lt = ce & 0xff;
            tc.rt = ce >> 8;
            if (a.cmacro) {
              const int m = a.cmacro[pid];
              tc.p0 = a.mpin[m];
              tc.p1 = a.mpin[m + 1];
            }
          }
          // every wave has finished its reads of the last commit's pixels
          if (NW > 1) __syncthreads();
          footprint_add<NW>(g, T, cx, cy, w, h, -1, rho, standard);  // rip-up
          if (tech) {  // doCommitOrRipup (AdmmKernels.hpp:467-506): the edges at the current orientation
            const bool was = will_rotate(h, T.vdd[g.ly(cy)], gnd);
            edge_add_tile(g, T, cx, cy, w, h, was ? tc.rt : tc.lt, was ? tc.lt : tc.rt, -1);
          }
          __syncthreads();
          SUB_MARK(11);
          float cst[NC];
          eval_candidates<MODE, NC>(a, g, T, disp_coeff, thr, cx, cy, w, h, code, gnd, pox, poy, d0, dir, cst, ucode,
                                    vmask, tc);
          SUB_MARK(12);
          // Tree-min of AdmmKernels.hpp:847-858 over slots s = j*64 + lane: a
          // halving-stride tournament in which a tie keeps the lower slot.  Its
          // winner is the lexicographic minimum of (cost, bit-reverse8(slot)),
          // since the last comparison decides on bit 0 of the slot, the one
          // before on bit 1, ...  Costs are >= 0 or +inf, so their bit patterns
          // order like the values.
          unsigned long long best = ~0ull;
  #pragma unroll
          for (int j = 0; j < NC; j++) {
            unsigned slot = (unsigned)(d0 + j * kWave);
            unsigned rev = __brev(slot) >> 24;
            unsigned long long key = ((unsigned long long)__float_as_uint(cst[j]) << 32) | rev;
            best = key < best ? key : best;
          }
          // minimum over the wave by DPP steps (keys are unique, so the order
          // of the steps does not matter); lane 63 ends with it
          unsigned bhi = (unsigned)(best >> 32), blo = (unsigned)best;
          key_min_step<0xB1, 0xf>(bhi, blo);   // quad_perm [1,0,3,2]
          key_min_step<0x4E, 0xf>(bhi, blo);   // quad_perm [2,3,0,1]
          key_min_step<0x141, 0xf>(bhi, blo);  // row_half_mirror
          key_min_step<0x140, 0xf>(bhi, blo);  // row_mirror
          key_min_step<0x142, 0xa>(bhi, blo);  // row_bcast:15 into rows 1, 3
          key_min_step<0x143, 0xc>(bhi, blo);  // row_bcast:31 into rows 2, 3
          unsigned bc_bits = __builtin_amdgcn_readlane(bhi, 63);
          unsigned bc_lo = __builtin_amdgcn_readlane(blo, 63);
          if (NW > 1) {
            // minimum over the waves' minima
            if (lane == kWave - 1) {
              xmin[wave] = bhi;
              xmin[NW + wave] = blo;
            }
            __syncthreads();
  #pragma unroll
            for (int w2 = 0; w2 < NW; w2++) {
              const unsigned ohi = xmin[w2], olo = xmin[NW + w2];
              const bool take = ohi < bc_bits || (ohi == bc_bits && olo < bc_lo);
              bc_bits = take ? ohi : bc_bits;
              bc_lo = take ? olo : bc_lo;
            }
          }
          const int bi = (int)(bc_lo & 0xffu);
          const int bslot = (int)(__brev((unsigned)bi) >> 24);
          int nx = cx, ny = cy;
          if (__uint_as_float(bc_bits) != kInf) {
            int2 dd = c_dir_default[bslot];
            new_pos(a, dd, w, h, cx, cy, nx, ny);
          }
          if (NW == 1) __syncthreads();
          SUB_MARK(13);
          footprint_add<NW>(g, T, nx, ny, w, h, +1, rho, standard);  // commit
          if (tech) {
            const bool will = will_rotate(h, T.vdd[g.ly(ny)], gnd);
            edge_add_tile(g, T, nx, ny, w, h, will ? tc.rt : tc.lt, will ? tc.lt : tc.rt, +1);
          }
          if (tid == 0) cellrec[pid].x = rec_pos(nx, ny);
          __syncthreads();
          SUB_MARK(14);
          start = first + 1;
          if (start >= kWave) break;
          // a move changes usage only inside the old and new footprints
          if (lane >= start && t < count && MODE == MODE_PRIMAL && !pert) {
            bool hit_old = my_x < cx + w && cx < my_x + my_w && my_y < cy + h && cy < my_y + my_h;
            bool hit_new = my_x < nx + w && nx < my_x + my_w && my_y < ny + h && ny < my_y + my_h;
            if (hit_old || hit_new) need = skip_test();
          }
          SUB_MARK(15);
        }
      }
    }
    PHASE_MARK(2);
  }

  // -- write back pixels; primal mode fuses the dual update and the overflow
  unsigned long long ovf = 0;
  {
    // 8 (one wave) or 4 (four waves) pixels per thread at a time, their LDS
    // reads issued together; the primal step also sets the next iteration's
    // activity flags
    constexpr int U = NW > 1 ? 4 : 8;
    int lb = tid / g.xh, la = tid - lb * g.xh;
    for (int i0 = 0; i0 < npix; i0 += U * NT) {
      int px[U], py[U], li[U];
#pragma unroll
      for (int q = 0; q < U; q++) {
        px[q] = g.gx(la);
        py[q] = g.gy(lb);
        li[q] = T.at(la, lb);
        if (i0 + q * NT + tid >= npix || py[q] < 0) px[q] = -1;
        la += step_a; lb += step_b;
        if (la >= g.xh) { la -= g.xh; lb++; }
      }
      uint32_t uc[U];
      float lm[U];
#pragma unroll
      for (int q = 0; q < U; q++) {
        const int l = px[q] >= 0 ? li[q] : 0;
        uc[q] = T.uc[l];
        lm[q] = T.lam[l];
      }
#pragma unroll
      for (int q = 0; q < U; q++) {
        if (px[q] < 0) continue;
        const unsigned p = pidx(a, px[q], py[q]);
        const int ui = (int)(uc[q] & 0xffffu);
        a.usage[p] = ui;
        if (MODE == MODE_PRIMAL) {
          a.dual[p] = dual_step(ui, lm[q], rho, step, neg);
          ovf += (unsigned long long)max(ui - 1, 0);
          if (ui > 1) flag_next(a, iter, px[q], py[q]);
        }
      }
    }
  }
  if (NW == 1 && a.tech) {  // the tile's edge maps back to the live maps
    for (int i = tid; i < npix; i += NT) {
      const int lb = i / g.xh, la = i - lb * g.xh;
      const int xx = g.gx(la), yy = g.gy(lb);
      if (xx < 0 || yy < 0) continue;
      const int li = T.at(la, lb);
      const unsigned p = pidx(a, xx, yy);
      a.edgeL[p] = T.eL[li];
      a.edgeR[p] = T.eR[li];
    }
  }
  if (MODE == MODE_PRIMAL) {
    for (int o = 32; o > 0; o >>= 1) ovf += __shfl_down(ovf, o);
    if (lane == 0 && ovf) atomicAdd(&ctl->ovf_acc, ovf);
    if (lazy) stamp_partition(a, g, iter, tid, NT);
  }
  PHASE_MARK(3);
  SUB_FLUSH();
  PHASE_END();
}
template <int MODE, int LDS_SORT, int NW>
__global__ __launch_bounds__(NW * 64) void k_partition(Args a, int npad_lo, int npad_hi) {
  partition_body<MODE, LDS_SORT, NW>(a, blockIdx.x, npad_lo, npad_hi);
}


// Dual update of the pixels of partitions that take no primal step
// (dual_update_functor_with_step_size, AdmmKernels.hpp:26-42).  Runs only
// while the activity flags are in use, beside the partition kernels: the two
// touch disjoint pixels.  A pixel of usage 1 has a zero
// increment, so its dual is read only for other usages.  Four consecutive
// pixels of a row per thread, rows in blockIdx.y.  The activity flags of the
// next iteration are set by the primal step itself.
__device__ __forceinline__ void dual_pass_body(const Args& a, int bx, int by) {
  const Ctl* ctl = a.ctl;
  if (ctl->done) return;
  const int iter = ctl->iter;
  const bool adaptive_now = a.use_adaptive && ctl->exceptional;
  if (iter < a.iter_perturbation_on || ctl->wrong_row != 0 || adaptive_now) return;
  // lazy duals: the increments of the last kCatchWin iterations in LDS
  __shared__ float s_catch[kCatchWin];
  if (a.lazy) load_catch_window(a, s_catch, iter + 1 - kCatchWin, iter, threadIdx.x, blockDim.x);
  const int y = by;
  const int x0 = 4 * (bx * blockDim.x + threadIdx.x);
  if (x0 >= a.x_size) return;
  const unsigned p0 = pidx(a, x0, y);
  const bool full = x0 + 4 <= a.x_size && (a.x_size & 3) == 0;
  int u[4];
  if (full) {
    const int4 v = *reinterpret_cast<const int4*>(a.usage + p0);
    u[0] = v.x; u[1] = v.y; u[2] = v.z; u[3] = v.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) u[q] = x0 + q < a.x_size ? a.usage[p0 + q] : 1;
  }
  const int tf = triple_fold(a, iter);
  const int rk = a.rowmap[tf * a.y_size + y];
  const int* cm = a.colmap + tf * a.x_size;
  const int* fl = a.act_flags + (size_t)(iter & 1) * a.K;
  bool upd[4];
  bool any = false;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    upd[q] = x0 + q < a.x_size && u[q] != 1 && fl[cm[x0 + q] * a.nrow + rk] == 0;
    any |= upd[q];
  }
  if (!any) return;
  const float rho = ctl->rho, step = ctl->step, neg = -1.0f / rho;
  // lazy duals (the periodic catch-up pass): every update since the
  // segment's stamp, this iteration's included; the stamps are set by
  // k_stamp_inactive after the pass
  int from[4];
#pragma unroll
  for (int q = 0; q < 4; q++) from[q] = (a.lazy && upd[q]) ? a.seg_stamp[seg_of(a, x0 + q, y)] + 1 : iter;
  auto upd1 = [&](int q, float l) {
    return a.lazy ? dual_catch_up(a, u[q], l, from[q], iter, s_catch, iter + 1 - kCatchWin)
                  : dual_step(u[q], l, rho, step, neg);
  };
  // a vector store only when all four pixels are updated here: a pixel of a
  // partition the primal step processes may share the group, and that step
  // runs at the same time
  if (full && upd[0] && upd[1] && upd[2] && upd[3]) {
    float4 d = *reinterpret_cast<const float4*>(a.dual + p0);
    d.x = upd1(0, d.x);
    d.y = upd1(1, d.y);
    d.z = upd1(2, d.z);
    d.w = upd1(3, d.w);
    *reinterpret_cast<float4*>(a.dual + p0) = d;
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (upd[q]) a.dual[p0 + q] = upd1(q, a.dual[p0 + q]);
  }
}

// Lazy duals, after the catch-up pass: the segments of partitions that took
// no primal step hold this iteration's update now.  One thread per segment
// (a segment lies in one partition).
__device__ __forceinline__ void stamp_inactive_body(const Args& a, int i0, int stride) {
  const Ctl* ctl = a.ctl;
  if (ctl->done || !a.lazy) return;
  const int iter = ctl->iter;
  if (iter < a.iter_perturbation_on || ctl->wrong_row != 0) return;
  const int tf = triple_fold(a, iter);
  const int* fl = a.act_flags + (size_t)(iter & 1) * a.K;
  const int ns = a.nxs * a.nys;
  for (int sg = i0; sg < ns; sg += stride) {
    const int sy = sg / a.nxs, sx = sg - sy * a.nxs;
    const int k = a.colmap[tf * a.x_size + a.xseg0[sx]] * a.nrow + a.rowmap[tf * a.y_size + a.yseg0[sy]];
    if (fl[k] == 0) a.seg_stamp[sg] = iter;
  }
}
__global__ __launch_bounds__(256) void k_stamp_inactive(Args a) {
  stamp_inactive_body(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Lazy duals, end of the solve (before the final refine changes usage):
// every price brought up to the last executed iteration.  One thread per
// pixel, rows in blockIdx.y.
__global__ __launch_bounds__(256) void k_dual_final(Args a) {
  const Ctl* ctl = a.ctl;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (!a.lazy || x >= a.x_size) return;
  const int last = ctl->iter - 1;
  const int s = a.seg_stamp[seg_of(a, x, y)];
  if (s >= last) return;
  const unsigned p = pidx(a, x, y);
  const float rho = ctl->rho;
  a.dual[p] = dual_catch_up(a, a.usage[p], a.dual[p], s + 1, last);
}
__global__ __launch_bounds__(256) void k_dual_pass(Args a) { dual_pass_body(a, blockIdx.x, blockIdx.y); }

// Before the first iteration: the activity flags of iteration 0 (any pixel
// with usage > 1 under its fold offset).  Four pixels per thread.
__global__ __launch_bounds__(256) void k_classify_init(Args a) {
  const Ctl* ctl = a.ctl;
  const int y = blockIdx.y;
  const int x0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (x0 >= a.x_size) return;
  const int iter = ctl->iter;
  const int tf = triple_fold(a, iter);
  const int rk = a.rowmap[tf * a.y_size + y];
  int32_t* fc = a.act_flags + (size_t)(iter & 1) * a.K;
  for (int q = 0; q < 4 && x0 + q < a.x_size; q++)
    if (a.usage[pidx(a, x0 + q, y)] > 1) fc[a.colmap[tf * a.x_size + x0 + q] * a.nrow + rk] = 1;
}

// Cell records <-> SoA positions (records hold positions during the solve)
__global__ void k_pack_rec(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int4 r;
  r.x = rec_pos(a.lx[c], a.ly[c]);
  r.y = a.w[c] | (a.h[c] << 12) | (a.gnd[c] << 18) | (a.code[c] << 19);
  r.z = a.ox[c];
  r.w = a.oy[c];
  reinterpret_cast<int4*>(a.cellrec)[c] = r;
  // moves never leave a cell on a wrong-parity row (such candidates cost
  // +inf), so this count only decreases during the solve
  if (!correct_row(a.h[c], a.gnd[c], a.vdd_up[a.ly[c]])) atomicAdd(&a.ctl->wrong_row, 1);
}

__global__ void k_unpack_pos(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int p = a.cellrec[4 * c];
  a.lx[c] = rec_x(p);
  a.ly[c] = rec_y(p);
}

// examineIteration + detectSlowConvergence + convergence test
// (CudaAdmmLegalizer.cu:305-318, 799-840, 995-1009).  Single thread.
__device__ __forceinline__ void finish_body(const Args& a) {
  Ctl* ctl = a.ctl;
  if (ctl->done) return;
  int iter = ctl->iter;
  unsigned long long ovf64 = ctl->ovf_acc;
  ctl->ovf_acc = 0;
  ctl->last_ovf = ovf64;
  int ovf = (int)ovf64;
  if (ovf < ctl->min_ovf) ctl->min_ovf = ovf;
  a.trace[iter] = ovf;
  // stagnation window (deque of the last `stagnation_window` overflows)
  int wn = min(a.stagnation_window, 64);
  if (ctl->ovf_window_n < wn) {
    ctl->ovf_window[ctl->ovf_window_n++] = ovf;
  } else {
    for (int i = 1; i < wn; i++) ctl->ovf_window[i - 1] = ctl->ovf_window[i];
    if (wn > 0) ctl->ovf_window[wn - 1] = ovf;
  }
  bool det = false;
  if (iter < a.min_iter_adaptive) det = false;
  else if (ctl->min_ovf == 0) det = false;
  else if (ctl->last_slow > 0 && iter < ctl->last_slow + a.cooltime_adaptive) det = false;
  else {
    bool by_ratio = float(ovf) / float(ctl->min_ovf) >= a.ovf_ratio_adaptive;
    bool by_stag = true;
    if (ctl->ovf_window_n == a.stagnation_window) {
      int f = ctl->ovf_window[0];
      for (int i = 0; i < ctl->ovf_window_n; i++)
        if (ctl->ovf_window[i] != f) { by_stag = false; break; }
    }
    det = by_ratio || by_stag;
    if (det) ctl->last_slow = iter;
  }
  ctl->exceptional = det ? 1 : 0;
  if (ovf <= a.terminate_thr) {
    ctl->converged = 1;
    ctl->conv_iter = iter;
    ctl->done = 1;
  }
  ctl->iter = iter + 1;
  if (ctl->iter >= a.max_iter) ctl->done = 1;
}
__global__ void k_finish(Args a) { finish_body(a); }

// ---------------------------------------------------------------------------
// adaptive search: extractCongestedCellsKernel + projectionAwareSearchKernel
// (AdmmKernels.hpp:947-1174), a single 1024-thread workgroup on global memory
// ---------------------------------------------------------------------------
__device__ __forceinline__ void extract_congested_body(const Args& a, int bx) {
  Ctl* ctl = a.ctl;
  if (ctl->done || !a.use_adaptive || !ctl->exceptional) return;
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int p0 = a.cellrec[4 * c];
  int cx = p0 & 0xffff, cy = (int)((unsigned)p0 >> 16);
  int s = 0;
  for (int x = cx; x < cx + a.w[c]; x++)
    for (int y = cy; y < cy + a.h[c]; y++) s += max(a.usage[pidx(a, x, y)] - 1, 0);
  a.ovf_flag[c] = s == 0 ? 0 : 1;
}
__global__ void k_extract_congested(Args a) { extract_congested_body(a, blockIdx.x); }

__device__ void proj_pos(const Args& a, int d, int w, int h, int cx, int cy, int& nx, int& ny) {
  int2 dd = c_dir_wide[d];
  int eps = (h % 2 == 0) ? 2 : 1;
  nx = cx + dd.x;
  ny = cy + eps * dd.y;
  int rid = -1;
  bool inval = false;
  for (int x = nx; x < nx + w && !inval; x++)
    for (int y = ny; y < ny + h; y++) {
      if (x >= a.x_size || y >= a.y_size || x < 0 || y < 0) { inval = true; break; }
      const unsigned p = pidx(a, x, y);
      if (a.pcode[p] == 0) { inval = true; rid = a.pregion[p]; break; }
    }
  if (rid != -1) {
    if (abs(dd.x) >= abs(dd.y)) nx = dd.x > 0 ? a.rxmax[rid] + 1 : a.rxmin[rid] - 1;
    else ny = dd.y > 0 ? a.rymax[rid] + 1 : a.rymin[rid] - 1;
  }
  nx = max(0, min(a.x_size - w, nx));
  ny = max(0, min(a.y_size - h, ny));
}

__device__ __forceinline__ void adaptive_body(const Args& a) {
  Ctl* ctl = a.ctl;
  if (ctl->done || !a.use_adaptive || !ctl->exceptional) return;
  __shared__ float s_cost[kNumDirWidePadded];
  __shared__ int s_idx[kNumDirWidePadded];
  __shared__ int s_scan[1024];
  __shared__ int s_n;
  __shared__ int s_ovf;
  const int t = threadIdx.x;
  // stream compaction in ascending cell id (thrust::copy_if)
  if (t == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < a.N; base += 1024) {
    int c = base + t;
    int f = (c < a.N) ? a.ovf_flag[c] : 0;
    s_scan[t] = f;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int v = (t >= o) ? s_scan[t - o] : 0;
      __syncthreads();
      s_scan[t] += v;
      __syncthreads();
    }
    if (f) a.ovf_list[s_n + s_scan[t] - 1] = c;
    __syncthreads();
    if (t == 0) s_n += s_scan[1023];
    __syncthreads();
  }
  const int n = s_n;
  const float rho = ctl->rho;
  for (int i = 0; i < n; i++) {
    int c = a.ovf_list[i];
    int p0 = a.cellrec[4 * c];
    int cx = p0 & 0xffff, cy = (int)((unsigned)p0 >> 16), w = a.w[c], h = a.h[c];
    if (t == 0) {
      int s = 0;
      for (int x = cx; x < cx + w; x++)
        for (int y = cy; y < cy + h; y++) s += max(a.usage[pidx(a, x, y)] - 1, 0);
      s_ovf = s;
    }
    __syncthreads();
    if (s_ovf == 0) { __syncthreads(); continue; }
    for (int q = t; q < w * h; q += 1024) a.usage[pidx(a, cx + q / h, cy + q % h)] -= 1;
    if (a.tech && t < h) {  // the edges follow the move (doCommitOrRipup); no pin / edge cost here
      const int ce = a.cedge[c];
      const bool r0 = will_rotate(h, a.vdd_up[cy], a.gnd[c]);
      const int lt = r0 ? ce >> 8 : ce & 0xff, rt = r0 ? ce & 0xff : ce >> 8;
      if (lt) a.edgeL[pidx(a, cx, cy + t)] += edge_delta(lt, -1);
      if (rt) a.edgeR[pidx(a, cx + w - 1, cy + t)] += edge_delta(rt, -1);
    }
    __syncthreads();
    float disp_coeff = 1.0f;
    if (a.flag_iccad17) disp_coeff = float(a.N) / float(a.k_height_num[h - 1]);
    int code = a.code[c], gnd = a.gnd[c];
    for (int d = t; d < kNumDirWidePadded; d += 1024) {
      float cost = kInf;
      if (d < kNumDirWide) {
        int nx, ny;
        proj_pos(a, d, w, h, cx, cy, nx, ny);
        float dc = disp_cost(a, disp_coeff, disp_threshold(a), nx, ny, a.ox[c], a.oy[c]);
        float pc;
        if (nx + w > a.x_size || ny + h > a.y_size || nx < 0 || ny < 0) pc = kInf;
        else if (!correct_row(h, gnd, a.vdd_up[ny])) pc = kInf;
        else {
          pc = 0.0f;
          for (int x = nx; x < nx + w && pc != kInf; x++)
            for (int y = ny; y < ny + h; y++) {
              const unsigned p = pidx(a, x, y);
              if ((int)a.pcode[p] != code) { pc = kInf; break; }
              float u = float(a.usage[p]);
              float l = a.dual[p];
              float m1 = u - 1.0f;
              float m = l * m1;
              float b = l / rho;
              pc += rho * fmax_ref(m + b, 0.0f);
            }
        }
        float qd = dc / a.coeff_neglect_disp;
        cost = qd + pc;
      }
      s_cost[d] = cost;
      s_idx[d] = d;
    }
    __syncthreads();
    for (int s = kNumDirWidePadded / 2; s > 0; s >>= 1) {
      for (int q = t; q < s; q += 1024)
        if (s_cost[q] > s_cost[q + s]) { s_cost[q] = s_cost[q + s]; s_idx[q] = s_idx[q + s]; }
      __syncthreads();
    }
    int nx = cx, ny = cy;
    if (s_cost[0] != kInf) proj_pos(a, s_idx[0], w, h, cx, cy, nx, ny);
    __syncthreads();
    for (int q = t; q < w * h; q += 1024) a.usage[pidx(a, nx + q / h, ny + q % h)] += 1;
    if (a.tech && t < h) {
      const int ce = a.cedge[c];
      const bool r1 = will_rotate(h, a.vdd_up[ny], a.gnd[c]);
      const int lt = r1 ? ce >> 8 : ce & 0xff, rt = r1 ? ce & 0xff : ce >> 8;
      if (lt) a.edgeL[pidx(a, nx, ny + t)] += edge_delta(lt, 1);
      if (rt) a.edgeR[pidx(a, nx + w - 1, ny + t)] += edge_delta(rt, 1);
    }
    if (t == 0) a.cellrec[4 * c] = (nx & 0xffff) | (ny << 16);
    __syncthreads();
  }
  if (t == 0) ctl->adaptive_calls++;
}
__global__ __launch_bounds__(1024) void k_adaptive(Args a) { adaptive_body(a); }

// ---------------------------------------------------------------------------
// post-process and checks
// ---------------------------------------------------------------------------

// -- technology maps and checks ------------------------------------------------

__global__ __launch_bounds__(256) void k_edge_clear(Args a) {
  const size_t P = (size_t)a.x_size * a.y_size;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
    a.edgeL[p] = kEdgeBias;
    a.edgeR[p] = kEdgeBias;
  }
}

// computePixelEdgeMapKernel (DatabaseKernels.hpp:105-164): one thread per
// cell; the orientation is "rotated" whenever the row's rail differs from
// the cell's, for every height, as that kernel takes it
__global__ __launch_bounds__(256) void k_edge_init(Args a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  const int x = a.lx[c], y = a.ly[c], w = a.w[c], h = a.h[c];
  if (x < 0 || y < 0 || x + w > a.x_size || y + h > a.y_size) return;
  const int ce = a.cedge[c];
  const bool rot = a.vdd_up[y] != a.gnd[c];
  const int lt = rot ? ce >> 8 : ce & 0xff, rt = rot ? ce & 0xff : ce >> 8;
  for (int yy = y; yy < y + h; yy++) {
    if (lt) atomicAdd(&a.edgeL[pidx(a, x, yy)], edge_delta(lt, 1));
    if (rt) atomicAdd(&a.edgeR[pidx(a, x + w - 1, yy)], edge_delta(rt, 1));
  }
}

// the edge maps as the next primal or refine step begins
__global__ __launch_bounds__(256) void k_edge_snap(Args a) {
  const size_t P = (size_t)a.x_size * a.y_size;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
    a.edgeLs[p] = a.edgeL[p];
    a.edgeRs[p] = a.edgeR[p];
  }
}

// countPinVioKernel + countEdgeSpacingVioKernel (DatabaseKernels.hpp:286-466)
// on the live maps: a short and an access violation count once per cell;
// edge_vio sums the neighbours' edges within spacing on both sides (each
// pair is seen from both cells)
__global__ __launch_bounds__(256) void k_tech_check(Args a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long shrt = 0, acc = 0, edges = 0;
  if (c < a.N) {
    const int x = a.lx[c], y = a.ly[c], w = a.w[c], h = a.h[c];
    if (x >= 0 && y >= 0 && x + w <= a.x_size && y + h <= a.y_size) {
      const bool rot = a.vdd_up[y] != a.gnd[c];
      if (a.cmacro) {
        const int m = a.cmacro[c];
        for (int i = a.mpin[m]; i < a.mpin[m + 1]; i++) {
          const int4 pn = a.pins[i];
          int bx = pn.x / h, by = pn.x % h, ex = pn.y / h, ey = pn.y % h;
          if (rot) {
            const int ox = bx, oy = by;
            bx = w - 1 - ex; by = h - 1 - ey;
            ex = w - 1 - ox; ey = h - 1 - oy;
          }
          for (int xx = x + min(bx, ex); xx <= x + max(bx, ex); xx++)
            for (int yy = y + min(by, ey); yy <= y + max(by, ey); yy++) {
              const int l = a.pmetal[pidx(a, xx, yy)];
              if (l > 0 && l == pn.z) shrt = 1;
              else if (l > 0 && l == pn.z + 1) acc = 1;
            }
        }
      }
      const int ce = a.cedge[c];
      const int lt = rot ? ce >> 8 : ce & 0xff, rt = rot ? ce & 0xff : ce >> 8;
      int n = 0;
      auto window = [&](const uint32_t* map, int type, int x0, int x1) {
        for (int xx = max(x0, 0); xx < min(x1, a.x_size); xx++)
          for (int yy = y; yy < y + h; yy++) n += edge_count(map[pidx(a, xx, yy)], type);
      };
      if (lt > 0) {
        window(a.edgeR, 1, x - a.rules[lt + 1], x);
        window(a.edgeR, 2, x - a.rules[lt + 2], x);
      }
      if (rt > 0) {
        window(a.edgeL, 1, x + w, x + w + a.rules[rt + 1]);
        window(a.edgeL, 2, x + w, x + w + a.rules[rt + 2]);
      }
      edges = (unsigned long long)(long long)n;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    shrt += __shfl_down(shrt, o);
    acc += __shfl_down(acc, o);
    edges += __shfl_down(edges, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (shrt) atomicAdd(&a.ctl->pin_short, shrt);
    if (acc) atomicAdd(&a.ctl->pin_access, acc);
    if (edges) atomicAdd(&a.ctl->edge_vio, edges);
  }
}

// rotateKernel (PostProcessKernel.hpp:6-26)
__global__ void k_rotate(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  if ((a.h[c] & 1) == 1) a.need_rot[c] = (a.gnd[c] == a.vdd_up[a.ly[c]]) ? 0 : 1;
}

// overflow of the recounted usage (CudaDatabase::computeOverflow)
__global__ void k_overflow(Args a) {
  size_t P = (size_t)a.x_size * a.y_size;
  unsigned long long s = 0;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x)
    s += (unsigned long long)max(a.usage[p] - 1, 0);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(&a.ctl->check_ovf, s);
}

// countRowOrientVioKernel (DatabaseKernels.hpp:468-504) + displacement sums,
// and the strict legality counts: per cell, whether its footprint leaves the
// grid, covers a pixel that is no valid site, covers a site of another fence
// group, and (even heights) whether its bottom row has the wrong rail.
__global__ void k_cell_stats(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long vio = 0, disp = 0, dmax = 0;
  int site = 0, fence = 0, core = 0, rail = 0;
  if (c < a.N) {
    const int lx = a.lx[c], ly = a.ly[c], w = a.w[c], hh = a.h[c];
    core = (lx < 0 || ly < 0 || lx + w > a.x_size || ly + hh > a.y_size) ? 1 : 0;
    int vdd = (ly >= 0 && ly < a.y_size) ? a.vdd_up[ly] : -1;
    int gnd = a.gnd[c], rot = a.need_rot[c];
    int v = 0;
    if ((hh & 1) == 1) {
      if (rot == 0 && gnd != vdd) v = 1;
      else if (rot == 1 && gnd == vdd) v = 1;
    } else if (gnd != vdd) {
      v = 1;
      rail = 1;
    }
    vio = v;
    const int code = a.code[c];
    const int x0 = max(lx, 0), x1 = min(lx + w, a.x_size);
    const int y0 = max(ly, 0), y1 = min(ly + hh, a.y_size);
    for (int y = y0; y < y1; y++)
      for (int x = x0; x < x1; x++) {
        const int pc = a.pcode[pidx(a, x, y)];
        if (pc == 0) site = 1;
        else if (pc != code) fence = 1;
      }
    int dx = lx * a.site_width + a.core_lx, dy = ly * a.row_height + a.core_ly;
    disp = (unsigned long long)(abs(dx - a.ox[c]) + abs(dy - a.oy[c]));
    dmax = disp;
  }
  for (int o = 32; o > 0; o >>= 1) {
    vio += __shfl_down(vio, o);
    disp += __shfl_down(disp, o);
    unsigned long long m = __shfl_down(dmax, o);
    dmax = m > dmax ? m : dmax;
    site += __shfl_down(site, o);
    fence += __shfl_down(fence, o);
    core += __shfl_down(core, o);
    rail += __shfl_down(rail, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (vio) atomicAdd(&a.ctl->row_vio, vio);
    if (disp) atomicAdd(&a.ctl->disp_sum, disp);
    atomicMax(&a.ctl->disp_max, dmax);
    if (site) atomicAdd(&a.ctl->site_vio, (unsigned long long)site);
    if (fence) atomicAdd(&a.ctl->fence_vio, (unsigned long long)fence);
    if (core) atomicAdd(&a.ctl->core_vio, (unsigned long long)core);
    if (rail) atomicAdd(&a.ctl->rail_vio, (unsigned long long)rail);
  }
}

// ---------------------------------------------------------------------------
// Batched launches of the request server.  One launch runs one step of an
// ADMM iteration for every request slot on the device: `tab` holds the
// slots' Args and block b of the launch does the work of block map[b].y of
// slot map[b].x's own (solo) launch, through the same body as the solo
// kernel, so a request computes exactly what it computes alone.  Every body
// returns at once for a slot whose Ctl says done, so a slot that finishes in
// the middle of a batch of iterations is left as it is until the host
// retires it.
// ---------------------------------------------------------------------------
#define LG_BATCH_SLOT()                                                        \
  const int2 m_ = map[blockIdx.x];                                             \
  const Args& a = tab[m_.x]

__global__ __launch_bounds__(256) void kb_bucket(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  bucket_body(a, m_.y);
}

template <bool LARGE, int NT>
__global__ __launch_bounds__(NT) void kb_scan(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  scan_body<LARGE, NT>(a);
}

__global__ void kb_extract_congested(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  if (!a.use_adaptive) return;
  extract_congested_body(a, m_.y);
}

__global__ __launch_bounds__(1024) void kb_adaptive(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  if (!a.use_adaptive) return;
  adaptive_body(a);
}

// reg_filter: 1 = only slots whose sorting algorithm is a bitonic network
// (sorted in registers up to kNpadSmall slots), 0 = only the others, -1 = all
template <int MODE, int LDS_SORT, int NW>
__global__ __launch_bounds__(NW * 64) void kb_partition(const Args* __restrict__ tab, const int2* __restrict__ map,
                                                        int npad_lo, int npad_hi, int reg_filter) {
  LG_BATCH_SLOT();
  const int reg = (a.sort_algo == LG_SORT_M_BITONIC || a.sort_algo == LG_SORT_S_BITONIC) ? 1 : 0;
  if (reg_filter >= 0 && reg != reg_filter) return;
  partition_body<MODE, LDS_SORT, NW>(a, m_.y, npad_lo, npad_hi);
}

// The large class of the primal step, solo: gridDim.x workgroups, workgroup
// g taking the entries g, g + gridDim.x, ... of the class-1 list.
template <int MODE, int LDS_SORT, int NW>
__global__ __launch_bounds__(NW * 64) void k_partition_large(Args a, int npad_lo, int npad_hi) {
  if (a.ctl->done) return;
  const int n = a.ctl->n_work[1];
  for (int j = blockIdx.x; j < n; j += gridDim.x) {
    partition_body<MODE, LDS_SORT, NW>(a, j, npad_lo, npad_hi);
    __syncthreads();  // the LDS tile is staged anew for the next entry
  }
}

// The large class of the primal step, batched: min(K, kLargeGroups)
// workgroups per slot (map {slot, g}), workgroup g taking the entries g,
// g + G, g + 2G, ... of the slot's class-1 list, whose length the scan and
// the four-wave launch fixed before this launch.  The list is short, so a
// workgroup per partition of every slot would mostly exit at once; the
// partitions of one step are independent, so the order changes nothing.
template <int MODE, int LDS_SORT, int NW>
__global__ __launch_bounds__(NW * 64) void kb_partition_large(const Args* __restrict__ tab,
                                                              const int2* __restrict__ map, int npad_lo,
                                                              int npad_hi) {
  LG_BATCH_SLOT();
  if (a.ctl->done) return;
  const int G = min(a.K, kLargeGroups);
  const int n = a.ctl->n_work[1];
  for (int j = m_.y; j < n; j += G) {
    partition_body<MODE, LDS_SORT, NW>(a, j, npad_lo, npad_hi);
    __syncthreads();  // the LDS tile is staged anew for the next entry
  }
}

// local block l of the solo launch's (ceil(x_size / 1024), y_size) grid
__global__ __launch_bounds__(256) void kb_dual_pass(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  const int gx = (a.x_size + 1023) / 1024;
  const int by = m_.y / gx;
  dual_pass_body(a, m_.y - by * gx, by);
}

// one block per slot, looping over the slot's segments
__global__ __launch_bounds__(256) void kb_stamp_inactive(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  stamp_inactive_body(a, threadIdx.x, blockDim.x);
}

__global__ void kb_finish(const Args* __restrict__ tab, const int2* __restrict__ map) {
  LG_BATCH_SLOT();
  finish_body(a);
}

#undef LG_BATCH_SLOT

}  // namespace lgdev
