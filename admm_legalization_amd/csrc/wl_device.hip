// Weighted-average wirelength gradient and the target function's
// preconditioned total gradient for gfx950: C ABI of include/skyline_wl.h.
// Replaces skyplace::WireLengthGradient (wirelength_gradient/WireLengthGradient.cu)
// and jacobianPrecondition (problem_instance/ProblemInstanceKernel.hpp:74-100).
//
// Layout in HBM (one handle):
//   nets   net_start[num_nets + 1] int32, net_w[num_nets] float
//   pins   pin2cell int32, pin_off float2 (x, y), pin_xy float2 (current
//          centres), pin_slot int32 (the pin's position in the cell-major
//          order below, -1 for a pin of no movable cell); pins numbered net
//          by net
//   cells  cell_start[num_cells + 1]; cell_g float2 [pins of movable cells]:
//          the gradient terms of the last evaluation, cell by cell, each
//          cell's pins in ascending pin order (the order built on the host)
//          cxy float2 [num_cells]: the cell centres side by side
//   chunks chunk_net[2 * num_chunks]: [first net, end net) of each maximal
//          run of consecutive nets of at most kSmallDeg pins, cut at
//          kChunkPins pins or kChunkNets nets
//   big    ids of the nets above kSmallDeg pins
//
// One evaluation = k_wl_pack (the centres side by side) + k_wl_chunk (a
// workgroup per chunk: its pins staged in LDS with coalesced loads, extrema
// and sums by a thread per net, exponentials and terms by a thread per pin)
// + k_wl_net_big (a wave per larger net) writing each term to its cell-major
// slot, then k_wl_cell (a thread per cell) summing its contiguous run.  The
// reference scatters the pin terms with float atomicAdd (addPinGrad,
// WireLengthGradient.cu:162-185); every sum here runs in a fixed order, so
// repeated evaluations are bit-identical.  HBM traffic per evaluation, with
// the pin update fused: pin2cell 4 + offsets 8 + cell centre gather 8 +
// pin_xy write 8 + slot 4 + term write 8 + term read 8 = 48 B per pin, 8 B
// per net (start, weight), 8 B per cell (two outputs) + 4 (CSR start).  The
// centre gather and the term write land on random cells: each is a line of
// the cache per pin, which bounds the kernel below the HBM rate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/skyline_wl.h"

namespace {

constexpr int kSmallDeg = 64;      // nets up to this many pins go to the chunks
constexpr int kChunkPins = 512;    // pins staged per chunk workgroup
constexpr int kChunkNets = 256;    // nets per chunk (a byte indexes them)
constexpr int kChunkThreads = 256;
constexpr int kBigWaves = 4;       // nets per k_wl_net_big workgroup (a wave each)

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define WL_TRY(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(WL_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct WlArgs {
  int num_cells, num_nets, num_pins, num_big, num_chunks;
  const int* net_start;
  const float* net_w;
  const int* pin2cell;
  const float2* pin_off;
  float2* pin_xy;
  const int* pin_slot;
  float2* cell_g;
  const int* cell_start;
  float2* cxy;
  const int* chunk_net;
  const int* big;
  double* hpwl;       // [1] result of wl_compute_hpwl
  double* hpwl_part;  // [kHpwlBlocks] per-block partial sums
};

// Pin centre: from its movable cell's centre (FROM_CELLS), else the stored one
// (updatePinCoordinateKernel, WireLengthGradient.cu:187-211).
template <bool FROM_CELLS>
__device__ __forceinline__ float2 pin_at(const WlArgs& a, int p, const float2* cxy) {
  if (FROM_CELLS) {
    const int c = a.pin2cell[p];
    if (c >= 0) {
      const float2 o = a.pin_off[p];
      const float2 q = cxy[c];
      const float2 v = make_float2(q.x + o.x, q.y + o.y);
      a.pin_xy[p] = v;
      return v;
    }
  }
  return a.pin_xy[p];
}

// computePinGrad (WireLengthGradient.cu:108-160), one axis.
__device__ __forceinline__ float pin_term(float x, float g_inv, float ap, float am, float bp, float bm, float cp,
                                          float cm) {
  const float m = ((1.0f - x * g_inv) * bm + g_inv * cm) * am / bm / bm;
  const float q = ((1.0f + x * g_inv) * bp - g_inv * cp) * ap / bp / bp;
  return m - q;
}

// A workgroup per chunk of small nets: computeMinMax (:14-45),
// computeABCKernel (:47-106) and computePinGrad (:108-160).  The chunk's pins
// are contiguous; they are read once, with coalesced loads, into LDS.  The
// extrema and the sums run by a thread per net over its pins in pin order;
// the four exponentials and the term run by a thread per pin.
template <bool FROM_CELLS>
__global__ __launch_bounds__(kChunkThreads) void k_wl_chunk(WlArgs a, float g_inv, const float2* cxy) {
  __shared__ float2 v[kChunkPins];      // pin centres
  __shared__ float4 ex[kChunkPins];     // exp terms: a+ x, a+ y, a- x, a- y
  __shared__ uint8_t pnet[kChunkPins];  // the pin's net in the chunk
  __shared__ float4 mc[kChunkNets];     // extrema (max x, max y, min x, min y), then c sums
  __shared__ float4 sb[kChunkNets];     // b sums (b+ x, b+ y, b- x, b- y)
  const int n0 = a.chunk_net[2 * blockIdx.x], n1 = a.chunk_net[2 * blockIdx.x + 1];
  const int p0 = a.net_start[n0];
  const int np = a.net_start[n1] - p0, nn = n1 - n0;
  const int t = threadIdx.x;
  for (int i = t; i < np; i += kChunkThreads) v[i] = pin_at<FROM_CELLS>(a, p0 + i, cxy);
  __syncthreads();
  for (int j = t; j < nn; j += kChunkThreads) {
    const int s = a.net_start[n0 + j] - p0, e = a.net_start[n0 + j + 1] - p0;
    float mxx = 0.0f, mxy = 0.0f, mnx = FLT_MAX, mny = FLT_MAX;
    for (int i = s; i < e; i++) {
      const float2 q = v[i];
      mxx = fmaxf(q.x, mxx);
      mnx = fminf(q.x, mnx);
      mxy = fmaxf(q.y, mxy);
      mny = fminf(q.y, mny);
      pnet[i] = (uint8_t)j;
    }
    mc[j] = make_float4(mxx, mxy, mnx, mny);
  }
  __syncthreads();
  for (int i = t; i < np; i += kChunkThreads) {
    const float4 m = mc[pnet[i]];
    const float2 q = v[i];
    ex[i] = make_float4(expf((q.x - m.x) * g_inv), expf((q.y - m.y) * g_inv), expf(-(q.x - m.z) * g_inv),
                        expf(-(q.y - m.w) * g_inv));
  }
  __syncthreads();
  for (int j = t; j < nn; j += kChunkThreads) {
    const int s = a.net_start[n0 + j] - p0, e = a.net_start[n0 + j + 1] - p0;
    float bpx = 0, bpy = 0, bmx = 0, bmy = 0, cpx = 0, cpy = 0, cmx = 0, cmy = 0;
    for (int i = s; i < e; i++) {
      const float2 q = v[i];
      const float4 x = ex[i];
      bpx += x.x; bpy += x.y; bmx += x.z; bmy += x.w;
      cpx += q.x * x.x; cpy += q.y * x.y; cmx += q.x * x.z; cmy += q.y * x.w;
    }
    sb[j] = make_float4(bpx, bpy, bmx, bmy);
    mc[j] = make_float4(cpx, cpy, cmx, cmy);
  }
  __syncthreads();
  for (int i = t; i < np; i += kChunkThreads) {
    const int slot = a.pin_slot[p0 + i];
    if (slot < 0) continue;
    const int j = pnet[i];
    const float2 q = v[i];
    const float4 x = ex[i], b = sb[j], c = mc[j];
    const float w = a.net_w[n0 + j];
    const float gx = pin_term(q.x, g_inv, x.x, x.z, b.x, b.z, c.x, c.z);
    const float gy = pin_term(q.y, g_inv, x.y, x.w, b.y, b.w, c.y, c.w);
    a.cell_g[slot] = make_float2(gx * w, gy * w);
  }
}

// The cell centres side by side (one gather per pin instead of two).
__global__ __launch_bounds__(256) void k_wl_pack(int n, const float* cx, const float* cy, float2* cxy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) cxy[c] = make_float2(cx[c], cy[c]);
}

template <typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T* sh) {
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); i++) r = op(r, sh[i]);
  return r;
}

template <typename Op>
__device__ __forceinline__ float wave_reduce(float v, Op op) {
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  return v;
}

// A wave per net of more than kSmallDeg pins (same arithmetic; each lane
// takes every 64th pin and the per-net values are reduced across the wave).
template <bool FROM_CELLS>
__global__ __launch_bounds__(kBigWaves * 64) void k_wl_net_big(WlArgs a, float g_inv, const float2* cxy) {
  const int wv = blockIdx.x * kBigWaves + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (wv >= a.num_big) return;
  const int n = a.big[wv];
  const int s = a.net_start[n], e = a.net_start[n + 1];
  float mxx = 0.0f, mxy = 0.0f, mnx = FLT_MAX, mny = FLT_MAX;
  for (int p = s + lane; p < e; p += 64) {
    const float2 v = pin_at<FROM_CELLS>(a, p, cxy);
    mxx = fmaxf(v.x, mxx);
    mnx = fminf(v.x, mnx);
    mxy = fmaxf(v.y, mxy);
    mny = fminf(v.y, mny);
  }
  auto fmx = [](float x, float y) { return fmaxf(x, y); };
  auto fmn = [](float x, float y) { return fminf(x, y); };
  auto add = [](float x, float y) { return x + y; };
  mxx = wave_reduce(mxx, fmx);
  mnx = wave_reduce(mnx, fmn);
  mxy = wave_reduce(mxy, fmx);
  mny = wave_reduce(mny, fmn);
  float bpx = 0, bpy = 0, bmx = 0, bmy = 0, cpx = 0, cpy = 0, cmx = 0, cmy = 0;
  for (int p = s + lane; p < e; p += 64) {
    const float2 v = a.pin_xy[p];
    const float apx = expf((v.x - mxx) * g_inv), apy = expf((v.y - mxy) * g_inv);
    const float amx = expf(-(v.x - mnx) * g_inv), amy = expf(-(v.y - mny) * g_inv);
    bpx += apx; bpy += apy; bmx += amx; bmy += amy;
    cpx += v.x * apx; cpy += v.y * apy; cmx += v.x * amx; cmy += v.y * amy;
  }
  bpx = wave_reduce(bpx, add); bpy = wave_reduce(bpy, add);
  bmx = wave_reduce(bmx, add); bmy = wave_reduce(bmy, add);
  cpx = wave_reduce(cpx, add); cpy = wave_reduce(cpy, add);
  cmx = wave_reduce(cmx, add); cmy = wave_reduce(cmy, add);
  const float w = a.net_w[n];
  for (int p = s + lane; p < e; p += 64) {
    const int slot = a.pin_slot[p];
    if (slot < 0) continue;
    const float2 v = a.pin_xy[p];
    const float apx = expf((v.x - mxx) * g_inv), apy = expf((v.y - mxy) * g_inv);
    const float amx = expf(-(v.x - mnx) * g_inv), amy = expf(-(v.y - mny) * g_inv);
    float gx = pin_term(v.x, g_inv, apx, amx, bpx, bmx, cpx, cmx);
    float gy = pin_term(v.y, g_inv, apy, amy, bpy, bmy, cpy, cmy);
    a.cell_g[slot] = make_float2(gx * w, gy * w);
  }
}

// addPinGrad (WireLengthGradient.cu:162-185) as a gather: a thread per cell
// sums the terms of its pins (a contiguous run of cell_g) in ascending pin
// order.
__global__ __launch_bounds__(256) void k_wl_cell(WlArgs a, float* gx, float* gy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.num_cells) return;
  float sx = 0.0f, sy = 0.0f;
  const int e = a.cell_start[c + 1];
  for (int i = a.cell_start[c]; i < e; i++) {
    const float2 g = a.cell_g[i];
    sx += g.x;
    sy += g.y;
  }
  gx[c] = sx;
  gy[c] = sy;
}

__global__ __launch_bounds__(256) void k_wl_update(WlArgs a, const float* cx, const float* cy) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.num_pins) return;
  const int c = a.pin2cell[p];
  if (c >= 0) {
    const float2 o = a.pin_off[p];
    a.pin_xy[p] = make_float2(cx[c] + o.x, cy[c] + o.y);
  }
}

// computeNetBBox + computeHPWL (WireLengthGradient.cu:213-235, 383-401):
// (max - min) per net and axis with the reference's initial values, summed in
// FP64: per-block partials over a fixed grid, then k_wl_hpwl_sum adds them in
// block order (the same bits on every call).
constexpr int kHpwlBlocks = 1024;

__global__ __launch_bounds__(256) void k_wl_hpwl(WlArgs a, double* part) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.num_nets; n += gridDim.x * blockDim.x) {
    float mxx = 0.0f, mxy = 0.0f, mnx = FLT_MAX, mny = FLT_MAX;
    for (int p = a.net_start[n]; p < a.net_start[n + 1]; p++) {
      const float2 q = a.pin_xy[p];
      mxx = fmaxf(q.x, mxx);
      mnx = fminf(q.x, mnx);
      mxy = fmaxf(q.y, mxy);
      mny = fminf(q.y, mny);
    }
    v += (double)(mxx - mnx) + (double)(mxy - mny);
  }
  v = block_reduce(v, [](double x, double y) { return x + y; }, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(64) void k_wl_hpwl_sum(const double* part, int nparts, double* out) {
  __shared__ double sh[64];
  double s = 0.0;
  for (int k = threadIdx.x; k < nparts; k += 64) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 64; k++) t += sh[k];
    *out = t;
  }
}

// jacobianPrecondition (ProblemInstanceKernel.hpp:74-100)
__global__ __launch_bounds__(256) void k_precond(int n, float lambda, float min_precond, const float* num_pins,
                                                 const float* dens_precond, const float* wgx, const float* wgy,
                                                 const float* dgx, const float* dgy, float* gx, float* gy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int wp = (int)num_pins[i];
  float pre = (float)wp + lambda * dens_precond[i];
  pre = fmaxf(pre, min_precond);
  gx[i] = (wgx[i] + lambda * dgx[i]) / pre;
  gy[i] = (wgy[i] + lambda * dgy[i]) / pre;
}

}  // namespace

struct WlHandle {
  WlArgs a{};
  hipStream_t stream = nullptr;
  std::vector<void*> bufs;
};

namespace {

template <typename T>
int up(WlHandle* h, T** dst, const T* src, size_t n) {
  T* p = nullptr;
  WL_TRY(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
  h->bufs.push_back(p);
  if (src && n) WL_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
  *dst = p;
  return WL_OK;
}

hipStream_t pick(WlHandle* h, void* s) { return s ? (hipStream_t)s : h->stream; }

template <bool FROM_CELLS>
int launch_grad(WlHandle* h, const float* cx, const float* cy, float g_inv, float* gx, float* gy, hipStream_t s) {
  const WlArgs& a = h->a;
  if (FROM_CELLS && a.num_cells > 0)
    hipLaunchKernelGGL(k_wl_pack, dim3((a.num_cells + 255) / 256), dim3(256), 0, s, a.num_cells, cx, cy, a.cxy);
  if (a.num_chunks > 0)
    hipLaunchKernelGGL((k_wl_chunk<FROM_CELLS>), dim3(a.num_chunks), dim3(kChunkThreads), 0, s, a, g_inv, a.cxy);
  if (a.num_big > 0)
    hipLaunchKernelGGL((k_wl_net_big<FROM_CELLS>), dim3((a.num_big + kBigWaves - 1) / kBigWaves),
                       dim3(kBigWaves * 64), 0, s, a, g_inv, a.cxy);
  if (a.num_cells > 0)
    hipLaunchKernelGGL(k_wl_cell, dim3((a.num_cells + 255) / 256), dim3(256), 0, s, a, gx, gy);
  WL_TRY(hipGetLastError());
  return WL_OK;
}

}  // namespace

extern "C" {

const char* wl_last_error(void) { return g_err.c_str(); }

int wl_create(const WlNetlist* n, WlHandle** out) {
  if (!n || !out) return fail(WL_ERR_ARG, "null argument");
  if (n->num_cells < 0 || n->num_nets < 0 || n->num_pins < 0) return fail(WL_ERR_ARG, "negative size");
  if (!n->net_start || (n->num_pins > 0 && (!n->pin2cell || !n->pin_offset_x || !n->pin_offset_y ||
                                            !n->pin_x || !n->pin_y)))
    return fail(WL_ERR_ARG, "null netlist array");
  if (n->net_start[0] != 0 || n->net_start[n->num_nets] != n->num_pins)
    return fail(WL_ERR_ARG, "net_start must run from 0 to num_pins");
  // big nets, and chunks ([first net, end net) pairs): maximal runs of small
  // nets, cut at kChunkPins pins or kChunkNets nets
  std::vector<int> big, chunk;
  int c0 = -1, cpins = 0;  // the open chunk's first net (-1: none) and its pins
  for (int i = 0; i < n->num_nets; i++) {
    const int d = n->net_start[i + 1] - n->net_start[i];
    if (d < 0) return fail(WL_ERR_ARG, "net_start not ascending");
    const bool is_big = d > kSmallDeg;
    if (c0 >= 0 && (is_big || i - c0 == kChunkNets || cpins + d > kChunkPins)) {
      chunk.push_back(c0);
      chunk.push_back(i);
      c0 = -1;
    }
    if (is_big) {
      big.push_back(i);
      continue;
    }
    if (c0 < 0) {
      c0 = i;
      cpins = 0;
    }
    cpins += d;
  }
  if (c0 >= 0) {
    chunk.push_back(c0);
    chunk.push_back(n->num_nets);
  }
  std::vector<int> cstart(n->num_cells + 1, 0);
  for (int p = 0; p < n->num_pins; p++) {
    const int c = n->pin2cell[p];
    if (c >= n->num_cells) return fail(WL_ERR_ARG, "pin2cell beyond num_cells");
    if (c >= 0) cstart[c + 1]++;
  }
  for (int c = 0; c < n->num_cells; c++) cstart[c + 1] += cstart[c];
  // each pin's slot in the cell-major order (its cell's pins ascending)
  std::vector<int> slot(std::max(n->num_pins, 1), -1), fill(cstart.begin(), cstart.end() - 1);
  for (int p = 0; p < n->num_pins; p++)
    if (n->pin2cell[p] >= 0) slot[p] = fill[n->pin2cell[p]]++;
  std::vector<float2> off(n->num_pins), xy(n->num_pins);
  for (int p = 0; p < n->num_pins; p++) {
    off[p] = make_float2(n->pin_offset_x[p], n->pin_offset_y[p]);
    xy[p] = make_float2(n->pin_x[p], n->pin_y[p]);
  }
  std::vector<float> w(std::max(n->num_nets, 1), 1.0f);
  if (n->net_weight) std::copy(n->net_weight, n->net_weight + n->num_nets, w.begin());

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WL_ERR_DEVICE, "no HIP device");
  WlHandle* h = new WlHandle();
  WlArgs& a = h->a;
  a.num_cells = n->num_cells;
  a.num_nets = n->num_nets;
  a.num_pins = n->num_pins;
  a.num_big = (int)big.size();
  a.num_chunks = (int)chunk.size() / 2;
  int st = 0;
  int *d_ns, *d_p2c, *d_slot, *d_cs, *d_ch, *d_big;
  float* d_w;
  float2 *d_off, *d_xy, *d_g, *d_cxy;
  double *d_h, *d_hp;
  if ((st = up(h, &d_ns, n->net_start, (size_t)n->num_nets + 1)) || (st = up(h, &d_w, w.data(), w.size())) ||
      (st = up(h, &d_p2c, n->pin2cell, (size_t)n->num_pins)) || (st = up(h, &d_off, off.data(), off.size())) ||
      (st = up(h, &d_xy, xy.data(), xy.size())) || (st = up(h, &d_slot, slot.data(), slot.size())) ||
      (st = up(h, &d_g, (const float2*)nullptr, (size_t)std::max(cstart[n->num_cells], 1))) ||
      (st = up(h, &d_cs, cstart.data(), cstart.size())) || (st = up(h, &d_ch, chunk.data(), chunk.size())) ||
      (st = up(h, &d_cxy, (const float2*)nullptr, (size_t)n->num_cells)) ||
      (st = up(h, &d_big, big.data(), big.size())) || (st = up(h, &d_h, (const double*)nullptr, 1)) ||
      (st = up(h, &d_hp, (const double*)nullptr, (size_t)kHpwlBlocks))) {
    wl_destroy(h);
    return st;
  }
  a.net_start = d_ns; a.net_w = d_w; a.pin2cell = d_p2c; a.pin_off = d_off; a.pin_xy = d_xy;
  a.pin_slot = d_slot; a.cell_g = d_g; a.cell_start = d_cs; a.cxy = d_cxy; a.chunk_net = d_ch; a.big = d_big;
  a.hpwl = d_h; a.hpwl_part = d_hp;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    wl_destroy(h);
    return fail(WL_ERR_DEVICE, "hipStreamCreate");
  }
  *out = h;
  return WL_OK;
}

int wl_update_pin_coordinates(WlHandle* h, const float* cx, const float* cy, void* stream) {
  if (!h || (h->a.num_pins > 0 && (!cx || !cy))) return fail(WL_ERR_ARG, "null argument");
  if (h->a.num_pins > 0)
    hipLaunchKernelGGL(k_wl_update, dim3((h->a.num_pins + 255) / 256), dim3(256), 0, pick(h, stream), h->a, cx, cy);
  WL_TRY(hipGetLastError());
  return WL_OK;
}

int wl_compute_grad(WlHandle* h, float g_inv, float* gx, float* gy, void* stream) {
  if (!h || (h->a.num_cells > 0 && (!gx || !gy))) return fail(WL_ERR_ARG, "null argument");
  return launch_grad<false>(h, nullptr, nullptr, g_inv, gx, gy, pick(h, stream));
}

int wl_compute_grad_at(WlHandle* h, const float* cx, const float* cy, float g_inv, float* gx, float* gy,
                       void* stream) {
  if (!h || (h->a.num_cells > 0 && (!gx || !gy || !cx || !cy))) return fail(WL_ERR_ARG, "null argument");
  return launch_grad<true>(h, cx, cy, g_inv, gx, gy, pick(h, stream));
}

int wl_compute_hpwl_async(WlHandle* h, double* d_out, void* stream) {
  if (!h || !d_out) return fail(WL_ERR_ARG, "null argument");
  hipStream_t s = pick(h, stream);
  const int nb = std::max(1, std::min(kHpwlBlocks, (h->a.num_nets + 255) / 256));
  hipLaunchKernelGGL(k_wl_hpwl, dim3(nb), dim3(256), 0, s, h->a, h->a.hpwl_part);
  hipLaunchKernelGGL(k_wl_hpwl_sum, dim3(1), dim3(64), 0, s, h->a.hpwl_part, nb, d_out);
  WL_TRY(hipGetLastError());
  return WL_OK;
}

int wl_compute_hpwl(WlHandle* h, double* out, void* stream) {
  if (!h || !out) return fail(WL_ERR_ARG, "null argument");
  hipStream_t s = pick(h, stream);
  int st = wl_compute_hpwl_async(h, h->a.hpwl, stream);
  if (st) return st;
  WL_TRY(hipMemcpyAsync(out, h->a.hpwl, sizeof(double), hipMemcpyDeviceToHost, s));
  WL_TRY(hipStreamSynchronize(s));
  return WL_OK;
}

int wl_get_pins(WlHandle* h, float* x, float* y) {
  if (!h) return fail(WL_ERR_ARG, "null handle");
  WL_TRY(hipStreamSynchronize(h->stream));
  WL_TRY(hipDeviceSynchronize());
  std::vector<float2> v(h->a.num_pins);
  if (!v.empty()) WL_TRY(hipMemcpy(v.data(), h->a.pin_xy, v.size() * sizeof(float2), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < v.size(); i++) {
    if (x) x[i] = v[i].x;
    if (y) y[i] = v[i].y;
  }
  return WL_OK;
}

int wl_destroy(WlHandle* h) {
  if (!h) return WL_OK;
  for (void* b : h->bufs) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return WL_OK;
}

int tf_precondition(int32_t n, float lambda, float min_precond, const float* num_pins, const float* dens_precond,
                    const float* wgx, const float* wgy, const float* dgx, const float* dgy, float* gx, float* gy,
                    void* stream) {
  if (n < 0) return fail(WL_ERR_ARG, "negative size");
  if (n == 0) return WL_OK;
  if (!num_pins || !dens_precond || !wgx || !wgy || !dgx || !dgy || !gx || !gy)
    return fail(WL_ERR_ARG, "null argument");
  hipLaunchKernelGGL(k_precond, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, lambda, min_precond,
                     num_pins, dens_precond, wgx, wgy, dgx, dgy, gx, gy);
  WL_TRY(hipGetLastError());
  return WL_OK;
}

}  // extern "C"
