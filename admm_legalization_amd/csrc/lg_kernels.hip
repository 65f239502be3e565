// HIP kernels of the ADMM negotiation legalizer for gfx950 (MI355X).
//
// Data layout in HBM (one legalization handle):
//   pixels, x-major (pixel_id = x * y_size + y):
//     usage  int32   cells covering the pixel
//     pcode  int16   valid ? fence_group + 2 : 0   (0 = not placeable)
//     dual   float   negotiation price (Lagrange multiplier)
//     region int32   region id (adaptive search projection only)
//   cells, SoA int32: lx, ly (grid), w, h, code (= group + 2), gnd, ox, oy (dbu)
//
// One iteration = k_assign -> k_scan -> k_scatter -> k_partition<PRIMAL>
// (sort + sequential cell moves + fused dual update / overflow) -> k_finish.
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

// std::max(a, b) of the reference's host-side semantics: (a < b) ? b : a
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

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
  int32_t* part_of;
  int32_t* cnt;
  int32_t* off;
  int32_t* fill;
  int32_t* grouped;
  int32_t* trace;
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
  Ctl* ctl;
};

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
    for (int y = y0; y < y1; y++) atomicAdd(&a.usage[(size_t)x * a.y_size + y], 1);
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
        s += a.usage[(size_t)x * a.y_size + y];
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
        float u = float(a.usage[(size_t)x * a.y_size + y]);
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

// assignPartitionKernel (AdmmKernels.hpp:122-162) + histogram
__global__ void k_assign(Args a) {
  const Ctl* ctl = a.ctl;
  if (ctl->done) return;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int tf = triple_fold(a, ctl->iter);
  int oxs = a.x_hint / 3 * tf, oys = a.y_hint / 3 * tf;
  int lxg = a.lx[c] + oxs, lyg = a.ly[c] + oys;
  int uxg = lxg + a.w[c] - 1, uyg = lyg + a.h[c] - 1;
  int c1 = (lxg / a.x_hint) % a.ncol, c2 = (uxg / a.x_hint) % a.ncol;
  int r1 = (lyg / a.y_hint) % a.nrow, r2 = (uyg / a.y_hint) % a.nrow;
  int p = (c1 == c2 && r1 == r2) ? c1 * a.nrow + r1 : a.K;
  a.part_of[c] = p;
  if (p < a.K) atomicAdd(&a.cnt[p], 1);
}

// Exclusive scan of the partition histogram, max count -> npad, dual step
// schedule (updateDualStepSize, CudaAdmmLegalizer.cu:293-294, 687-691).
// One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(Args a) {
  Ctl* ctl = a.ctl;
  if (ctl->done) return;
  __shared__ int s_part[1024];
  __shared__ int s_max[1024];
  const int K = a.K;
  const int t = threadIdx.x;
  // each thread scans a contiguous chunk
  int per = (K + 1023) / 1024;
  int b = t * per, e = min(K, b + per);
  int sum = 0, mx = 0;
  for (int i = b; i < e; i++) { sum += a.cnt[i]; mx = max(mx, a.cnt[i]); }
  s_part[t] = sum;
  s_max[t] = mx;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = (t >= o) ? s_part[t - o] : 0;
    int m = (t + o < 1024) ? s_max[t + o] : 0;
    __syncthreads();
    s_part[t] += v;
    s_max[t] = max(s_max[t], m);
    __syncthreads();
  }
  int run = s_part[t] - sum;
  for (int i = b; i < e; i++) {
    a.off[i] = run;
    run += a.cnt[i];
    a.fill[i] = 0;
    a.cnt[i] = 0;
  }
  if (t == 0) {
    int total = s_part[1023];
    a.off[K] = total;
    a.off[K + 1] = total;
    int m = s_max[0];
    int np = 1;
    while (np < m) np <<= 1;
    ctl->max_cnt = m;
    ctl->npad = np;
    if (np > kMaxNpad) { ctl->error = kErrPartitionTooDense; ctl->done = 1; }
    int it = ctl->iter;
    if (it % a.iter_update_dual_step == 0 && it > a.iter_threshold_dual_update) {
      float inc = a.coeff_update_dual_step * ctl->lambda0;
      ctl->step = ctl->step + inc;
    }
  }
}

__global__ void k_scatter(Args a) {
  if (a.ctl->done) return;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int p = a.part_of[c];
  if (p >= a.K) return;
  int slot = atomicAdd(&a.fill[p], 1);
  a.grouped[a.off[p] + slot] = c;
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

// Ascending sort of distinct keys (cell ids) -- restores ascending-id order
// within a partition after the unordered scatter.
__device__ void lds_sort_ids(int n, int* v) {
  const int lane = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < n; i += kWave) {
        int ij = i ^ j;
        if (ij > i) {
          int a0 = v[i], a1 = v[ij];
          bool up = (i & k) == 0;
          if (up ? (a0 > a1) : (a0 < a1)) { v[i] = a1; v[ij] = a0; }
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

// Partition geometry for the iteration (paintPixelsKernel colour rule,
// AdmmKernels.hpp:164-191, evaluated arithmetically instead of stored).
struct PartGeo {
  int oxs, oys, col, row, xh, yh, ncol, nrow;
  __device__ bool in_x(int x) const { return ((x + oxs) / xh) % ncol == col; }
  __device__ bool in_y(int y) const { return ((y + oys) / yh) % nrow == row; }
  __device__ int lx(int x) const { return (x + oxs) % xh; }
  __device__ int ly(int y) const { return (y + oys) % yh; }
  // global coordinate of local column la / row lb, or -1 if outside the grid.
  // Local index la is reached from X = col*xh + la or (col+ncol)*xh + la
  // (the wrap-around of the shifted grid); at most one lies in the grid.
  __device__ int gx(int la, int x_size) const {
    int x = col * xh + la - oxs;
    if (x >= 0 && x < x_size) return x;
    x += ncol * xh;
    return (x >= 0 && x < x_size) ? x : -1;
  }
  __device__ int gy(int lb, int y_size) const {
    int y = row * yh + lb - oys;
    if (y >= 0 && y < y_size) return y;
    y += nrow * yh;
    return (y >= 0 && y < y_size) ? y : -1;
  }
};

// Cached cell record in LDS (16 bytes)
struct CellRec {
  int pos;    // lx | ly << 16 (grid)
  int shape;  // w | h << 12 | gnd << 18 | code << 19
  int ox, oy; // dbu
};

__device__ __forceinline__ int rec_w(int s) { return s & 0xfff; }
__device__ __forceinline__ int rec_h(int s) { return (s >> 12) & 0x3f; }
__device__ __forceinline__ int rec_gnd(int s) { return (s >> 18) & 1; }
__device__ __forceinline__ int rec_code(int s) { return (s >> 19) & 0x1fff; }

// computeDispCost (AdmmKernels.hpp:508-530)
__device__ __forceinline__ float disp_cost(const Args& a, float disp_coeff, int nx, int ny, int ox, int oy) {
  int dx = nx * a.site_width + a.core_lx;
  int dy = ny * a.row_height + a.core_ly;
  int disp = abs(dx - ox) + abs(dy - oy);
  float ds = float(disp) / float(a.site_width);
  float thr = float(a.max_disp_in_row_height) * float(a.row_height);
  thr = thr / float(a.site_width);
  float ex = ds - thr;
  float p = disp_coeff * ds;
  float q = a.max_disp_coeff * fmax_ref(ex, 0.0f);
  return p + q;
}

// getNewCellPosition (AdmmKernels.hpp:540-560)
__device__ __forceinline__ void new_pos(const Args& a, int2 d, int w, int h, int cx, int cy, int& nx, int& ny) {
  int eps = (h % 2 == 0) ? 2 : 1;
  nx = cx + d.x;
  ny = cy + eps * d.y;
  nx = max(0, min(a.x_size - w, nx));
  ny = max(0, min(a.y_size - h, ny));
}

template <int MODE>
struct PartitionKernel;

enum { MODE_PRIMAL = 0, MODE_REFINE = 1 };

// computePenaltySum (AdmmKernels.hpp:333-391) with partition_exclusive = true,
// evaluated on the LDS copy of the partition's pixels.
template <int MODE>
__device__ __forceinline__ float penalty_lds(const Args& a, const PartGeo& g, const uint2* pix,
                                             const int* vdd_loc, float rho, bool standard, int nx, int ny,
                                             int w, int h, int code, int gnd) {
  if (nx + w > a.x_size || ny + h > a.y_size || nx < 0 || ny < 0) return kInf;
  // all footprint pixels must carry this partition's colour
  if (g.ncol > 1 && !(w <= g.xh && g.in_x(nx) && g.in_x(nx + w - 1))) return kInf;
  if (g.nrow > 1 && !(h <= g.yh && g.in_y(ny) && g.in_y(ny + h - 1))) return kInf;
  int ly0 = g.ly(ny);
  if (!correct_row(h, gnd, vdd_loc[ly0])) return kInf;
  float s = 0.0f;
  int lxi = g.lx(nx);
  for (int x = 0; x < w; x++) {
    int lyi = ly0;
    for (int y = 0; y < h; y++) {
      uint2 e = pix[lxi * g.yh + lyi];
      if ((int)(e.x >> 16) != code) return kInf;
      float u = float(e.x & 0xffffu);
      float l = __uint_as_float(e.y);
      if (MODE == MODE_REFINE) {
        if (u > 0.0f) return kInf;
      } else {
        float t;
        float m1 = u - 1.0f;
        float b = l / rho;
        if (standard) {
          t = rho * fmax_ref(m1 + b, 0.0f);
        } else {
          float m = l * m1;
          t = rho * fmax_ref(m + b, 0.0f);
        }
        s += t;
      }
      if (++lyi == g.yh) lyi = 0;
    }
    if (++lxi == g.xh) lxi = 0;
  }
  return MODE == MODE_REFINE ? 0.0f : s;
}

__device__ __forceinline__ void footprint_add(const PartGeo& g, uint2* pix, int x0, int y0, int w, int h,
                                              int delta) {
  const int lane = threadIdx.x;
  const int n = w * h;
  int lx0 = g.lx(x0), ly0 = g.ly(y0);
  for (int i = lane; i < n; i += kWave) {
    int lxi = lx0 + i / h, lyi = ly0 + i % h;
    if (lxi >= g.xh) lxi -= g.xh;
    if (lyi >= g.yh) lyi -= g.yh;
    uint2* e = &pix[lxi * g.yh + lyi];
    e->x = (uint32_t)((int)e->x + delta);
  }
}

// doPrimalUpdateKernel (AdmmKernels.hpp:637-897) / doFinalRefineKernel
// (AdmmKernels.hpp:1176-1416) for one partition, plus -- in primal mode -- the
// dual update (dual_update_functor_with_step_size, AdmmKernels.hpp:26-42) and
// the overflow count of the partition's pixels.
template <int MODE>
__global__ __launch_bounds__(64) void k_partition(Args a, int npad_lo, int npad_hi) {
  Ctl* ctl = a.ctl;
  int iter;
  if (MODE == MODE_PRIMAL) {
    if (ctl->done) return;
    if (a.use_adaptive && ctl->exceptional) {
      if (npad_lo != 0) return;  // only the small variant runs the dual-only pass
    } else {
      int np = ctl->npad;
      if (np < npad_lo || np > npad_hi) return;
    }
    iter = ctl->iter;
  } else {
    int np = ctl->npad;
    if (np < npad_lo || np > npad_hi) return;
    iter = ctl->conv_iter;
  }
  const bool dual_only = (MODE == MODE_PRIMAL) && a.use_adaptive && ctl->exceptional;
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  PartGeo g;
  int tf = triple_fold(a, iter);
  g.xh = a.x_hint; g.yh = a.y_hint; g.ncol = a.ncol; g.nrow = a.nrow;
  g.oxs = g.xh / 3 * tf; g.oys = g.yh / 3 * tf;
  g.col = k / a.nrow; g.row = k % a.nrow;
  const int npad = dual_only ? 1 : ctl->npad;

  extern __shared__ __align__(16) unsigned char lds_raw[];
  uint2* pix = reinterpret_cast<uint2*>(lds_raw);                       // [xh*yh]
  int* vdd_loc = reinterpret_cast<int*>(pix + g.xh * g.yh);             // [yh]
  CellRec* rec = reinterpret_cast<CellRec*>(vdd_loc + ((g.yh + 3) & ~3));  // [npad]
  int* keys = reinterpret_cast<int*>(rec + npad_hi);                    // [npad]
  int* vals = keys + npad_hi;                                           // [npad]
  int* aux = vals + npad_hi;                                            // [npad] (merge/stable sorts)

  // -- stage the partition's pixels: local (a_, b_) <-> global (x, y) -------
  const int npix = g.xh * g.yh;
  for (int i = lane; i < npix; i += kWave) {
    int la = i / g.yh, lb = i - la * g.yh;
    int xx = g.gx(la, a.x_size), yy = g.gy(lb, a.y_size);
    uint2 e;
    if (xx >= 0 && yy >= 0) {
      size_t p = (size_t)xx * a.y_size + yy;
      int u = a.usage[p];
      if (u > 0xffff) ctl->error = kErrUsageRange;
      e.x = ((uint32_t)(uint16_t)a.pcode[p] << 16) | (uint32_t)(u & 0xffff);
      e.y = __float_as_uint(a.dual[p]);
    } else {
      e.x = 0;  // code 0: never placeable
      e.y = 0;
    }
    pix[i] = e;
  }
  for (int b = lane; b < g.yh; b += kWave) {
    int yy = g.gy(b, a.y_size);
    vdd_loc[b] = yy >= 0 ? a.vdd_up[yy] : 0;
  }
  __syncthreads();

  const int off = a.off[k];
  const int count = dual_only ? 0 : a.off[k + 1] - off;
  const float rho = ctl->rho;

  if (!dual_only && count > 0) {
    // -- processing order -------------------------------------------------
    if (MODE == MODE_PRIMAL) {
      int na = 1;
      while (na < count) na <<= 1;
      for (int i = lane; i < na; i += kWave) vals[i] = (i < count) ? a.grouped[off + i] : INT_MAX;
      __syncthreads();
      lds_sort_ids(na, vals);  // ascending cell id = thrust::stable_sort_by_key order
      // computeKeyArrayForSortingKernel (SortingKernels.hpp:9-55)
      int cr = 0, co = 10, ca = 1;
      if (a.sort_policy == 0) { cr = 1; co = 0; ca = 0; }
      for (int i = lane; i < npad; i += kWave) {
        if (i < count) {
          int c = vals[i];
          int x0 = a.lx[c], y0 = a.ly[c], w = a.w[c], h = a.h[c];
          int ovf = 0;
          int lx0 = g.lx(x0), ly0 = g.ly(y0);
          for (int x = 0; x < w; x++) {
            int lxi = lx0 + x;
            if (lxi >= g.xh) lxi -= g.xh;
            for (int y = 0; y < h; y++) {
              int lyi = ly0 + y;
              if (lyi >= g.yh) lyi -= g.yh;
              int u = (int)(pix[lxi * g.yh + lyi].x & 0xffffu);
              ovf += max(u - 1, 0);
            }
          }
          keys[i] = cr * a.rand_prio[c] + co * ovf + ca * w * h;
        } else {
          keys[i] = INT_MIN;
          vals[i] = -1;
        }
      }
      __syncthreads();
      if (a.sort_algo == 0 || a.sort_algo == 1) lds_merge_desc(npad, keys, vals, aux, reinterpret_cast<int*>(rec));
      else if (a.sort_algo == 4) lds_stable_desc(npad, count, keys, vals, aux);
      else lds_bitonic_desc(npad, keys, vals);
      for (int i = lane; i < count; i += kWave) a.grouped[off + i] = vals[i];
      __syncthreads();
    } else {
      for (int i = lane; i < count; i += kWave) vals[i] = a.grouped[off + i];
      __syncthreads();
    }
    // -- cache the cells' records ----------------------------------------
    for (int i = lane; i < count; i += kWave) {
      int c = vals[i];
      CellRec r;
      r.pos = (a.lx[c] & 0xffff) | (a.ly[c] << 16);
      r.shape = a.w[c] | (a.h[c] << 12) | (a.gnd[c] << 18) | (a.code[c] << 19);
      r.ox = a.ox[c];
      r.oy = a.oy[c];
      rec[i] = r;
    }
    __syncthreads();

    const bool pert = iter < a.iter_perturbation_on;
    const bool standard = (MODE == MODE_PRIMAL) && a.standard_admm;
    // -- the sequential chain of cells ------------------------------------
    for (int t = 0; t < count; t++) {
      CellRec r = rec[t];
      const int cx = r.pos & 0xffff, cy = (int)((unsigned)r.pos >> 16);
      const int w = rec_w(r.shape), h = rec_h(r.shape), gnd = rec_gnd(r.shape), code = rec_code(r.shape);
      if (MODE == MODE_PRIMAL) {
        bool ok_row = correct_row(h, gnd, vdd_loc[g.ly(cy)]);
        // computeCellOverflow (AdmmKernels.hpp:393-406)
        int ov = 0;
        int lx0 = g.lx(cx), ly0 = g.ly(cy);
        for (int i = lane; i < w * h; i += kWave) {
          int lxi = lx0 + i / h, lyi = ly0 + i % h;
          if (lxi >= g.xh) lxi -= g.xh;
          if (lyi >= g.yh) lyi -= g.yh;
          int u = (int)(pix[lxi * g.yh + lyi].x & 0xffffu);
          ov += max(u - 1, 0);
        }
        ov = wave_sum(ov);
        if (!pert && ov == 0 && ok_row) continue;
      } else {
        if (h <= 1) continue;
      }
      float disp_coeff = 1.0f;
      if (a.flag_iccad17) disp_coeff = float(a.N) / float(a.k_height_num[h - 1]);
      footprint_add(g, pix, cx, cy, w, h, -1);  // rip-up
      __syncthreads();
      float cst[4];
      int idx[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int d = j * kWave + lane;
        idx[j] = d;
        if (d < kNumDirDefault) {
          int nx, ny;
          new_pos(a, c_dir_default[d], w, h, cx, cy, nx, ny);
          float dc = disp_cost(a, disp_coeff, nx, ny, r.ox, r.oy);
          float pc = penalty_lds<MODE>(a, g, pix, vdd_loc, rho, standard, nx, ny, w, h, code, gnd);
          cst[j] = dc + pc;
        } else {
          cst[j] = kInf;
        }
      }
      // tree-min of AdmmKernels.hpp:847-858: slot s = j*64 + lane
      if (cst[0] > cst[2]) { cst[0] = cst[2]; idx[0] = idx[2]; }
      if (cst[1] > cst[3]) { cst[1] = cst[3]; idx[1] = idx[3]; }
      if (cst[0] > cst[1]) { cst[0] = cst[1]; idx[0] = idx[1]; }
      float bc = cst[0];
      int bi = idx[0];
#pragma unroll
      for (int s = 32; s > 0; s >>= 1) {
        float oc = __shfl_down(bc, s);
        int oi = __shfl_down(bi, s);
        if (lane < s && bc > oc) { bc = oc; bi = oi; }
      }
      bc = wave_bcast0(bc);
      bi = wave_bcast0(bi);
      int nx = cx, ny = cy;
      if (bc != kInf) new_pos(a, c_dir_default[bi], w, h, cx, cy, nx, ny);
      __syncthreads();
      footprint_add(g, pix, nx, ny, w, h, +1);  // commit
      if (lane == 0) rec[t].pos = (nx & 0xffff) | (ny << 16);
      __syncthreads();
    }
    // -- write back cell positions ----------------------------------------
    for (int i = lane; i < count; i += kWave) {
      int c = vals[i];
      int pos = rec[i].pos;
      a.lx[c] = pos & 0xffff;
      a.ly[c] = (int)((unsigned)pos >> 16);
    }
  }

  // -- write back pixels; primal mode fuses the dual update and the overflow
  unsigned long long ovf = 0;
  const float step = ctl->step;
  const float neg = -1.0f / rho;
  for (int i = lane; i < npix; i += kWave) {
    int la = i / g.yh, lb = i - la * g.yh;
    int xx = g.gx(la, a.x_size), yy = g.gy(lb, a.y_size);
    if (xx < 0 || yy < 0) continue;
    size_t p = (size_t)xx * a.y_size + yy;
    uint2 e = pix[i];
    int ui = (int)(e.x & 0xffffu);
    a.usage[p] = ui;
    if (MODE == MODE_PRIMAL) {
      float u = float(ui);
      float gg = fmax_ref(u - 1.0f, neg);
      double q = (double)rho * 0.5;
      q = q * (double)gg;
      q = q * (double)gg;
      float delta = (float)((double)gg + q);
      float m = step * delta;
      float l = __uint_as_float(e.y) + m;
      a.dual[p] = fmax_ref(l, 0.0f);
      ovf += (unsigned long long)max(ui - 1, 0);
    }
  }
  if (MODE == MODE_PRIMAL) {
    for (int o = 32; o > 0; o >>= 1) ovf += __shfl_down(ovf, o);
    if (lane == 0 && ovf) atomicAdd(&ctl->ovf_acc, ovf);
  }
}

// examineIteration + detectSlowConvergence + convergence test
// (CudaAdmmLegalizer.cu:305-318, 799-840, 995-1009).  Single thread.
__global__ void k_finish(Args a) {
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

// ---------------------------------------------------------------------------
// adaptive search: extractCongestedCellsKernel + projectionAwareSearchKernel
// (AdmmKernels.hpp:947-1174), a single 1024-thread workgroup on global memory
// ---------------------------------------------------------------------------
__global__ void k_extract_congested(Args a) {
  Ctl* ctl = a.ctl;
  if (ctl->done || !a.use_adaptive || !ctl->exceptional) return;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  int s = 0;
  for (int x = a.lx[c]; x < a.lx[c] + a.w[c]; x++)
    for (int y = a.ly[c]; y < a.ly[c] + a.h[c]; y++) s += max(a.usage[(size_t)x * a.y_size + y] - 1, 0);
  a.ovf_flag[c] = s == 0 ? 0 : 1;
}

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
      size_t p = (size_t)x * a.y_size + y;
      if (a.pcode[p] == 0) { inval = true; rid = a.pregion[p]; break; }
    }
  if (rid != -1) {
    if (abs(dd.x) >= abs(dd.y)) nx = dd.x > 0 ? a.rxmax[rid] + 1 : a.rxmin[rid] - 1;
    else ny = dd.y > 0 ? a.rymax[rid] + 1 : a.rymin[rid] - 1;
  }
  nx = max(0, min(a.x_size - w, nx));
  ny = max(0, min(a.y_size - h, ny));
}

__global__ __launch_bounds__(1024) void k_adaptive(Args a) {
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
    int cx = a.lx[c], cy = a.ly[c], w = a.w[c], h = a.h[c];
    if (t == 0) {
      int s = 0;
      for (int x = cx; x < cx + w; x++)
        for (int y = cy; y < cy + h; y++) s += max(a.usage[(size_t)x * a.y_size + y] - 1, 0);
      s_ovf = s;
    }
    __syncthreads();
    if (s_ovf == 0) { __syncthreads(); continue; }
    for (int q = t; q < w * h; q += 1024) a.usage[(size_t)(cx + q / h) * a.y_size + cy + q % h] -= 1;
    __syncthreads();
    float disp_coeff = 1.0f;
    if (a.flag_iccad17) disp_coeff = float(a.N) / float(a.k_height_num[h - 1]);
    int code = a.code[c], gnd = a.gnd[c];
    for (int d = t; d < kNumDirWidePadded; d += 1024) {
      float cost = kInf;
      if (d < kNumDirWide) {
        int nx, ny;
        proj_pos(a, d, w, h, cx, cy, nx, ny);
        float dc = disp_cost(a, disp_coeff, nx, ny, a.ox[c], a.oy[c]);
        float pc;
        if (nx + w > a.x_size || ny + h > a.y_size || nx < 0 || ny < 0) pc = kInf;
        else if (!correct_row(h, gnd, a.vdd_up[ny])) pc = kInf;
        else {
          pc = 0.0f;
          for (int x = nx; x < nx + w && pc != kInf; x++)
            for (int y = ny; y < ny + h; y++) {
              size_t p = (size_t)x * a.y_size + y;
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
    for (int q = t; q < w * h; q += 1024) a.usage[(size_t)(nx + q / h) * a.y_size + ny + q % h] += 1;
    if (t == 0) { a.lx[c] = nx; a.ly[c] = ny; }
    __syncthreads();
  }
  if (t == 0) ctl->adaptive_calls++;
}

// ---------------------------------------------------------------------------
// post-process and checks
// ---------------------------------------------------------------------------

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

// countRowOrientVioKernel (DatabaseKernels.hpp:468-504) + displacement sums
__global__ void k_cell_stats(Args a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long vio = 0, disp = 0, dmax = 0;
  if (c < a.N) {
    int vdd = a.vdd_up[a.ly[c]];
    int hh = a.h[c], gnd = a.gnd[c], rot = a.need_rot[c];
    int v = 0;
    if ((hh & 1) == 1) {
      if (rot == 0 && gnd != vdd) v = 1;
      else if (rot == 1 && gnd == vdd) v = 1;
    } else if (gnd != vdd) v = 1;
    vio = v;
    int dx = a.lx[c] * a.site_width + a.core_lx, dy = a.ly[c] * a.row_height + a.core_ly;
    disp = (unsigned long long)(abs(dx - a.ox[c]) + abs(dy - a.oy[c]));
    dmax = disp;
  }
  for (int o = 32; o > 0; o >>= 1) {
    vio += __shfl_down(vio, o);
    disp += __shfl_down(disp, o);
    unsigned long long m = __shfl_down(dmax, o);
    dmax = m > dmax ? m : dmax;
  }
  if ((threadIdx.x & 63) == 0) {
    if (vio) atomicAdd(&a.ctl->row_vio, vio);
    if (disp) atomicAdd(&a.ctl->disp_sum, disp);
    atomicMax(&a.ctl->disp_max, dmax);
  }
}

}  // namespace lgdev
