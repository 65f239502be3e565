// Global placer for gfx950: C ABI of include/skyline_gp.h.
//
// Replaces skyplace::SkyPlace::run with SkyPlaceDB::init (SkyPlaceDB.cpp:
// 140-355, 551-662), InitialPlacer::doRandomInit (InitialPlacer.cpp:77-123),
// TargetFunction (TargetFunction.cu:17-299) and NesterovSolver / AdamSolver
// (NesterovSolver.cu:54-264, AdamSolver.cu:38-226).  The objective's two
// terms are this library's WA wirelength (wl_device.hip) and electrostatic
// density (dens_device.hip) evaluators, called through their C ABI on the
// placer's stream.
//
// Layout in HBM: every solver vector holds 2 * n floats, x of all movable
// cells then y (n = movable instances + fillers, movable instances first in
// instance order), as the reference's CudaVector<float> of 2 * numMovable.
// Per cell: real width / height (clipping), pin count (Jacobi
// preconditioner).  Scalars that steer the loop (Lipschitz sums, HPWL,
// overflow, gradient 1-norms) are reduced on the device in a fixed order
// (per-block partials, then one block) and read back in one 48-byte copy per
// objective evaluation; the loop's decisions (backtracking, lambda, gamma,
// convergence) are taken on the host as in the reference.
//
// Kernels and their HBM bytes per vector entry (2n entries per launch):
//   k_move_forward   reads cur, cur_pre, grad, size (16 B), writes 2 (8 B)
//   k_move_backward  reads var, grad, size (12 B), writes 4 B
//   k_adam_step      reads var, grad, m, v, size (20 B), writes var, m, v (12 B)
//   k_reduce         reads 2 or 4 vectors (8 or 16 B)
// all HBM-bound; the objective's kernels dominate the iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/skyline_density.h"
#include "../../include/skyline_gp.h"
#include "../../include/skyline_wl.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define GP_TRY(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(GP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kThreads = 256;
constexpr int kRedBlocks = 1024;

struct Die {
  float lx, ly, ux, uy;
};

// Values read back after each objective evaluation.
struct Scalars {
  double q0, q1;   // Lipschitz: sum (dx)^2, sum (dgrad)^2; initial lambda: |wl|_1, |dens|_1
  double hpwl;     // computeHPWL at the evaluated point
  float overflow;  // DensityGradient::overflow() at the evaluated point
  float pad;
  double pad2[2];
};

// getXInsideChip / getYInsideChip (ProblemInstanceKernel.hpp:7-33)
__device__ __forceinline__ float inside(float c, float size, float lo, float hi) {
  float r = c;
  if (c - size / 2 < lo) r = lo + size / 2;
  if (c + size / 2 > hi) r = hi - size / 2;
  return r;
}

__device__ __forceinline__ void axis_of(int i, int n, const float* w, const float* h, const Die& d, float& size,
                                        float& lo, float& hi) {
  if (i < n) {
    size = w[i];
    lo = d.lx;
    hi = d.ux;
  } else {
    size = h[i - n];
    lo = d.ly;
    hi = d.uy;
  }
}

// moveForwardKernel + clipToChipBoundary of both outputs (NesterovSolver.cu:
// 29-47, 145-172): x' = pre + step * grad, pre' = x' + coef * (x' - x), each
// then moved inside the die.
__global__ __launch_bounds__(kThreads) void k_move_forward(int n, float step, float coef, const float* cur,
                                                           const float* pre, const float* grad, float* nvar,
                                                           float* npre, const float* w, const float* h, Die d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  float size, lo, hi;
  axis_of(i, n, w, h, d, size, lo, hi);
  const float nx = pre[i] + step * grad[i];
  const float np = nx + coef * (nx - cur[i]);
  nvar[i] = inside(nx, size, lo, hi);
  npre[i] = inside(np, size, lo, hi);
}

// moveBackwardKernel + clip (NesterovSolver.cu:16-27, 174-194)
__global__ __launch_bounds__(kThreads) void k_move_backward(int n, float step, const float* var, const float* grad,
                                                            float* out, const float* w, const float* h, Die d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  float size, lo, hi;
  axis_of(i, n, w, h, d, size, lo, hi);
  out[i] = inside(var[i] - step * grad[i], size, lo, hi);
}

// One Adam iteration after the gradient (AdamSolver.cu:79-160): moments
// m' = b1 m - (1 - b1) g, v' = b2 v + (1 - b2) g^2, bias-corrected direction
// m' c1 / (sqrt(max(v' c2, 0)) + eps), x' = x - alpha * direction, clipped.
__global__ __launch_bounds__(kThreads) void k_adam_step(int n, float alpha, float b1, float b2, float c1, float c2,
                                                        float eps, const float* var, const float* grad,
                                                        const float* m, const float* v, float* nvar, float* nm,
                                                        float* nv, const float* w, const float* h, Die d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  float size, lo, hi;
  axis_of(i, n, w, h, d, size, lo, hi);
  const float g = grad[i];
  const float m1 = b1 * m[i] + (-(1.0f - b1)) * g;
  const float v1 = b2 * v[i] + (1.0f - b2) * g * g;
  const float bc1 = c1 * m1, bc2 = fmaxf(0.0f, c2 * v1);
  const float dir = bc1 / (sqrtf(bc2) + eps);
  nm[i] = m1;
  nv[i] = v1;
  nvar[i] = inside(var[i] - alpha * dir, size, lo, hi);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// doRandomInit (InitialPlacer.cpp:77-123) with a counter-based generator:
// a normal pair per cell by Box-Muller in FP64, fillers mirrored about the
// die centre (getMirrorX / getMirrorY, InitialPlacer.cpp:735-751), the centre
// stored in FP32 and moved inside the die (moveCellInsideLayout,
// SkyPlaceDB.cpp:738-770).
__global__ __launch_bounds__(kThreads) void k_init_place(int n, int n_mov, uint64_t seed, double mx, double my,
                                                         double sx, double sy, Die d, const float* w, const float* h,
                                                         float* var) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = splitmix64(seed + 2ull * (uint64_t)i), b = splitmix64(seed + 2ull * (uint64_t)i + 1ull);
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(b >> 11) * 0x1.0p-53;          // [0, 1)
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586 * u2;
  double x = mx + sx * (r * cos(t));
  double y = my + sy * (r * sin(t));
  if (i >= n_mov) {
    x = x <= mx ? (double)d.lx + (mx - x) : (double)d.ux - (x - mx);
    y = y <= my ? (double)d.ly + (my - y) : (double)d.uy - (y - my);
  }
  var[i] = inside((float)x, w[i], d.lx, d.ux);
  var[n + i] = inside((float)y, h[i], d.ly, d.uy);
}

// Deterministic two-quantity reduction over m entries: SQDIFF sums (a-b)^2
// and (c-d)^2 (computeDelta2Norm squared), ABS sums |a| and |c|
// (computeNorm1).  Per-block partials in block order, then k_collect.
template <bool SQDIFF>
__global__ __launch_bounds__(kThreads) void k_reduce(int m, const float* a, const float* b, const float* c,
                                                     const float* e, double* part) {
  __shared__ double sh[2][kThreads / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    if (SQDIFF) {
      const double d0 = (double)a[i] - (double)b[i], d1 = (double)c[i] - (double)e[i];
      s0 += d0 * d0;
      s1 += d1 * d1;
    } else {
      s0 += fabs((double)a[i]);
      s1 += fabs((double)c[i]);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][wv] = s0;
    sh[1][wv] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int k = 0; k < kThreads / 64; k++) {
      t0 += sh[0][k];
      t1 += sh[1][k];
    }
    part[2 * blockIdx.x] = t0;
    part[2 * blockIdx.x + 1] = t1;
  }
}

// Sums the partials in block order and gathers the evaluation's scalars.
__global__ __launch_bounds__(64) void k_collect(const double* part, int nblk, const double* hpwl, const float* ovf,
                                                Scalars* out) {
  __shared__ double sh[2][64];
  double s0 = 0.0, s1 = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 64) {
    s0 += part[2 * k];
    s1 += part[2 * k + 1];
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int k = 0; k < 64; k++) {
      t0 += sh[0][k];
      t1 += sh[1][k];
    }
    Scalars r{};
    r.q0 = t0;
    r.q1 = t1;
    r.hpwl = hpwl ? *hpwl : 0.0;
    r.overflow = ovf ? *ovf : 0.0f;
    *out = r;
  }
}

int blocks(long long m) { return (int)std::max<long long>(1, (m + kThreads - 1) / kThreads); }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct GpHandle {
  GpParams p{};
  int num_insts = 0;
  int n_mov = 0, n_fill = 0, n = 0;  // movable instances, fillers, all movable
  std::vector<int> mov_inst;         // movable index -> instance index
  std::vector<int32_t> inst_lx, inst_ly;
  std::vector<float> w, h;           // real sizes (host copy)
  Die die{};
  float filler_w = 0.f, filler_h = 0.f;
  int nbx = 0, nby = 0;
  hipStream_t s = nullptr;
  WlHandle* wl = nullptr;
  DensHandle* dens = nullptr;
  std::vector<void*> bufs;
  // device vectors [2n]
  float *cur_var = nullptr, *new_var = nullptr;
  float *cur_pre = nullptr, *new_pre = nullptr, *prv_pre = nullptr;
  float *cur_grad = nullptr, *new_grad = nullptr, *prv_grad = nullptr;
  float *wl_grad = nullptr, *dens_grad = nullptr;
  float *m1 = nullptr, *v1 = nullptr, *m2 = nullptr, *v2 = nullptr;  // Adam moments (cur, new)
  // device per cell [n]
  float *dw = nullptr, *dh = nullptr, *num_pins = nullptr;
  double* part = nullptr;
  double* hpwl_acc = nullptr;
  Scalars* sc_dev = nullptr;
  Scalars* sc_host = nullptr;  // pinned
  // TargetFunction / solver state
  float lambda = 0.f, gamma_inv = 0.f, hpwl = 0.f, prev_hpwl = 0.f, overflow = 0.f;
  float step = 0.f, cur_a = 1.f;
  float alpha = 0.f, beta1k = 0.f, beta2k = 0.f;
  int iter = 0;
  int evals = 0;
  bool solver_ready = false, done = false, converged = false;
  int last_bt = 0;
  double hpwl_initial = 0.0;
  double t_init = 0.0, t_solve = 0.0;
  GpTrace trace{};
};

namespace {

template <typename T>
int dalloc(GpHandle* h, T** p, size_t n) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return fail(GP_ERR_MEMORY, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes");
  h->bufs.push_back(q);
  *p = (T*)q;
  return GP_OK;
}

template <typename T>
int dupload(GpHandle* h, T** p, const T* src, size_t n) {
  int st = dalloc(h, p, n);
  if (st) return st;
  if (n) GP_TRY(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return GP_OK;
}

int wl_status(int st) { return st ? fail(GP_ERR_DEVICE, std::string("wirelength: ") + wl_last_error()) : GP_OK; }
int dens_status(int st) { return st ? fail(GP_ERR_DEVICE, std::string("density: ") + dens_last_error()) : GP_OK; }

// updatePointAndGetGrad (TargetFunction.cu:88-141) at var into grad with the
// given lambda, then the scalars: the Lipschitz sums of (ref_var - var) and
// (ref_grad - grad) when ref_var is given, HPWL and overflow at var.
int evaluate(GpHandle* h, const float* var, float* grad, float lambda, const float* ref_var, const float* ref_grad) {
  const int n = h->n;
  int st;
  if ((st = wl_status(wl_compute_grad_at(h->wl, var, var + n, h->gamma_inv, h->wl_grad, h->wl_grad + n, h->s))))
    return st;
  if ((st = dens_status(dens_compute_grad(h->dens, var, var + n, h->dens_grad, h->dens_grad + n, nullptr, h->s))))
    return st;
  if ((st = wl_status(tf_precondition(n, lambda, h->p.min_precond, h->num_pins, dens_get_preconditioner(h->dens),
                                      h->wl_grad, h->wl_grad + n, h->dens_grad, h->dens_grad + n, grad, grad + n,
                                      h->s))))
    return st;
  h->evals++;
  const int m = 2 * n;
  const int nb = std::min(kRedBlocks, blocks(m));
  if (ref_var)
    hipLaunchKernelGGL(k_reduce<true>, dim3(nb), dim3(kThreads), 0, h->s, m, ref_var, var, ref_grad, grad, h->part);
  if ((st = wl_status(wl_compute_hpwl_async(h->wl, h->hpwl_acc, h->s)))) return st;
  hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, h->s, h->part, ref_var ? nb : 0, h->hpwl_acc,
                     dens_get_overflow_device(h->dens), h->sc_dev);
  GP_TRY(hipGetLastError());
  GP_TRY(hipMemcpyAsync(h->sc_host, h->sc_dev, sizeof(Scalars), hipMemcpyDeviceToHost, h->s));
  GP_TRY(hipStreamSynchronize(h->s));
  return GP_OK;
}

// computeLipschitz (NesterovSolver.cu:196-212) from the reduced sums
float lipschitz(const Scalars& s, int m) {
  const double a = std::sqrt(s.q0 / (double)m), b = std::sqrt(s.q1 / (double)m);
  return (float)(a / b);
}

// updateGammaInv (TargetFunction.cu:241-256)
void update_gamma(GpHandle* h, float ovf) {
  float g;
  if (ovf > 1.0f)
    g = 0.1f;
  else if (ovf < 0.1f)
    g = 10.0f;
  else
    g = (float)(1.0 / std::pow(10.0, 20 * (ovf - 0.1) / 9.0 - 1.0));
  h->gamma_inv = g * h->p.init_gamma_inv_coef;
}

// updateLambda (TargetFunction.cu:258-274)
void update_lambda(GpHandle* h) {
  const float p = (h->hpwl - h->prev_hpwl) / h->p.reference_hpwl;
  float coef;
  if (p < 0)
    coef = h->p.max_phi_coef;
  else
    coef = std::max(h->p.min_phi_coef, (float)(h->p.max_phi_coef * std::pow(h->p.max_phi_coef, -p)));
  h->lambda *= coef;
}

// updateParameters (TargetFunction.cu:276-285) from the last evaluation
void update_parameters(GpHandle* h) {
  h->prev_hpwl = h->hpwl;
  h->hpwl = (float)h->sc_host->hpwl;
  h->overflow = h->sc_host->overflow;
  update_gamma(h, h->overflow);
  update_lambda(h);
}

// checkConvergence (TargetFunction.cu:287-299)
bool check_convergence(const GpHandle* h) {
  return h->overflow <= h->p.target_overflow && !(h->hpwl < h->prev_hpwl);
}

void record(GpHandle* h, int it, int bt) {
  GpTrace& t = h->trace;
  if (it < 0 || it >= t.cap) return;
  if (t.hpwl) t.hpwl[it] = h->hpwl;
  if (t.overflow) t.overflow[it] = h->overflow;
  if (t.lambda) t.lambda[it] = h->lambda;
  if (t.step) t.step[it] = h->p.solver == GP_SOLVER_ADAM ? h->alpha : h->step;
  if (t.backtracks) t.backtracks[it] = bt;
}

void progress(const GpHandle* h, int it, double secs) {
  if (h->p.verbose) std::printf("| %5d | %12.1f |  %-4.4f  | %8.2f |\n", it, h->hpwl, h->overflow, secs);
}

// getInitialGrad (TargetFunction.cu:177-239): density first (its overflow
// sets gamma), wirelength, the gradient preconditioned with lambda = 0, and
// lambda scaled by |wl grad|_1 / |density grad|_1.
int initial_grad(GpHandle* h, const float* var, float* grad) {
  const int n = h->n;
  int st;
  if ((st = dens_status(dens_compute_grad(h->dens, var, var + n, h->dens_grad, h->dens_grad + n, nullptr, h->s))))
    return st;
  float ovf = 0.f;
  GP_TRY(hipMemcpyAsync(&h->sc_host->overflow, dens_get_overflow_device(h->dens), sizeof(float),
                        hipMemcpyDeviceToHost, h->s));
  GP_TRY(hipStreamSynchronize(h->s));
  ovf = h->sc_host->overflow;
  h->overflow = ovf;
  update_gamma(h, ovf);
  if ((st = wl_status(wl_compute_grad_at(h->wl, var, var + n, h->gamma_inv, h->wl_grad, h->wl_grad + n, h->s))))
    return st;
  if ((st = wl_status(tf_precondition(n, 0.0f, h->p.min_precond, h->num_pins, dens_get_preconditioner(h->dens),
                                      h->wl_grad, h->wl_grad + n, h->dens_grad, h->dens_grad + n, grad, grad + n,
                                      h->s))))
    return st;
  h->evals++;
  const int m = 2 * n;
  const int nb = std::min(kRedBlocks, blocks(m));
  hipLaunchKernelGGL(k_reduce<false>, dim3(nb), dim3(kThreads), 0, h->s, m, h->wl_grad, nullptr, h->dens_grad,
                     nullptr, h->part);
  hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, h->s, h->part, nb, nullptr, nullptr, h->sc_dev);
  GP_TRY(hipGetLastError());
  GP_TRY(hipMemcpyAsync(h->sc_host, h->sc_dev, sizeof(Scalars), hipMemcpyDeviceToHost, h->s));
  GP_TRY(hipStreamSynchronize(h->s));
  const float wl_sum = (float)h->sc_host->q0, dens_sum = (float)h->sc_host->q1;
  if (!std::isfinite(wl_sum) || !std::isfinite(dens_sum) || !(dens_sum > 0.f))
    return fail(GP_ERR_ARG, "initial gradient sums are not finite (wirelength " + std::to_string(wl_sum) +
                                ", density " + std::to_string(dens_sum) + ")");
  h->lambda *= wl_sum / dens_sum;
  return GP_OK;
}

void swap_vec(float*& a, float*& b) { std::swap(a, b); }

// NesterovSolver::backTracking (NesterovSolver.cu:123-151).  Every trial of
// the reference's loop moves forward with the same step_length_ from the same
// point and evaluates the objective with the same lambda and gamma (none of
// them changes inside the loop), so its trials after the first repeat the
// first; this evaluator gives equal bits for equal inputs, so one trial
// yields the loop's result: the new step if it is at least 0.95 times the
// old, else the minimum step if below it, else the new step.
int backtrack(GpHandle* h, float coef, float& out_step, int& evals) {
  const int n = h->n;
  hipLaunchKernelGGL(k_move_forward, dim3(blocks(2LL * n)), dim3(kThreads), 0, h->s, n, h->step, coef, h->cur_var,
                     h->cur_pre, h->cur_grad, h->new_var, h->new_pre, h->dw, h->dh, h->die);
  int st = evaluate(h, h->new_pre, h->new_grad, h->lambda, h->cur_pre, h->cur_grad);
  if (st) return st;
  float ns = lipschitz(*h->sc_host, 2 * n);
  if (!(ns >= h->step * 0.95) && ns < h->p.min_step_length) ns = h->p.min_step_length;
  out_step = ns;
  evals = 1;
  return GP_OK;
}

int nesterov_iteration(GpHandle* h) {
  const float prev_a = h->cur_a;
  h->cur_a = (float)(0.5 * (1.0 + std::sqrt(4.0 * prev_a * prev_a + 1.0)));
  const float coef = (float)((prev_a - 1.0) / h->cur_a);
  float ns;
  int bt;
  int st = backtrack(h, coef, ns, bt);
  if (st) return st;
  h->step = ns;
  h->last_bt = bt;
  // updateOneIteration (NesterovSolver.cu:245-258)
  swap_vec(h->prv_pre, h->cur_pre);
  swap_vec(h->prv_grad, h->cur_grad);
  swap_vec(h->cur_pre, h->new_pre);
  swap_vec(h->cur_grad, h->new_grad);
  swap_vec(h->cur_var, h->new_var);
  return GP_OK;
}

int adam_iteration(GpHandle* h) {
  const int n = h->n;
  // gradient at the current point (the scalars serve updateParameters)
  int st = evaluate(h, h->cur_var, h->cur_grad, h->lambda, nullptr, nullptr);
  if (st) return st;
  const float c1 = (float)(1.0 / (1.0 - h->beta1k)), c2 = (float)(1.0 / (1.0 - h->beta2k));
  hipLaunchKernelGGL(k_adam_step, dim3(blocks(2LL * n)), dim3(kThreads), 0, h->s, n, h->alpha, h->p.adam_beta1,
                     h->p.adam_beta2, c1, c2, h->p.adam_epsilon, h->cur_var, h->cur_grad, h->m1, h->v1, h->new_var,
                     h->m2, h->v2, h->dw, h->dh, h->die);
  GP_TRY(hipGetLastError());
  swap_vec(h->cur_var, h->new_var);
  swap_vec(h->m1, h->m2);
  swap_vec(h->v1, h->v2);
  h->alpha *= 0.997f;
  h->beta1k *= h->p.adam_beta1;
  h->beta2k *= h->p.adam_beta2;
  h->last_bt = 1;
  return GP_OK;
}

// createFillers (SkyPlaceDB.cpp:551-662) in FP32 as the reference: the
// filler size is the mean width and height over the 10th..90th percentile
// range of the width-sorted list (the height list is left unsorted there),
// the count fills target density x white space minus the scaled movable
// area; centres from std::mt19937(0).
int create_fillers(const std::vector<float>& w, const std::vector<float>& hh, const Die& die, float sum_fixed,
                   float sum_scaled, float td, float& fw, float& fh, int& count, std::vector<float>& fcx,
                   std::vector<float>& fcy) {
  const int nm = (int)w.size();
  std::vector<float> dx(w), dy(hh);
  std::sort(dx.begin(), dx.end());
  int lo = (int)((float)dx.size() * 0.10);
  int hi = (int)((float)dx.size() * 0.90);
  if (lo == hi) {
    lo = 0;
    hi = nm;
  }
  float sx = 0.f, sy = 0.f;
  for (int i = lo; i < hi; i++) {
    sx += dx[i];
    sy += dy[i];
  }
  fw = sx / (float)(hi - lo);
  fh = sy / (float)(hi - lo);
  const float avg_area = fw * fh;
  const float die_area = (die.ux - die.lx) * (die.uy - die.ly);
  const float white = die_area - sum_fixed;
  const float filler_area = white * td - sum_scaled;
  if (filler_area < 0)
    return fail(GP_ERR_ARG, "filler area is negative: use a higher target density");
  count = (int)(filler_area / avg_area);
  std::mt19937 rnd(0);
  const int dlx = (int)die.lx, dly = (int)die.ly, dwid = (int)(die.ux - die.lx), dhgt = (int)(die.uy - die.ly);
  if (count > 0 && (dwid <= 0 || dhgt <= 0)) return fail(GP_ERR_ARG, "empty die");
  fcx.resize(count);
  fcy.resize(count);
  for (int i = 0; i < count; i++) {
    const auto rx = rnd();
    const auto ry = rnd();
    fcx[i] = (float)(rx % (decltype(rx))dwid + (decltype(rx))dlx);
    fcy[i] = (float)(ry % (decltype(ry))dhgt + (decltype(ry))dly);
  }
  return GP_OK;
}

int check_params(const GpParams* p) {
  if (p->max_opt_iter < 0 || p->max_backtrack_iter < 1) return fail(GP_ERR_ARG, "iteration limits");
  if (!(p->target_density > 0.f) || p->target_density > 1.f) return fail(GP_ERR_ARG, "target density not in (0, 1]");
  if (p->solver != GP_SOLVER_NESTEROV && p->solver != GP_SOLVER_ADAM) return fail(GP_ERR_ARG, "unknown solver");
  if (p->precision < 0 || p->precision > 2) return fail(GP_ERR_ARG, "unknown precision");
  if ((p->num_bin_x > 0) != (p->num_bin_y > 0) || p->num_bin_x < 0 || p->num_bin_y < 0)
    return fail(GP_ERR_ARG, "give both bin counts or neither");
  if (!(p->reference_hpwl > 0.f) || !(p->min_step_length > 0.f) || p->reference_hpwl_ratio < 0.f)
    return fail(GP_ERR_ARG, "bad hyper-parameter");
  if (!(p->scaling_factor > 0.f)) return fail(GP_ERR_ARG, "scaling factor must be positive");
  return GP_OK;
}

}  // namespace

extern "C" {

const char* gp_last_error(void) { return g_err.c_str(); }

void gp_default_params(GpParams* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->max_opt_iter = 2500;
  p->max_backtrack_iter = 10;
  p->init_lambda = 1e-4f;
  p->init_gamma_inv_coef = 0.01f;
  p->target_overflow = 0.07f;
  p->min_phi_coef = 0.95f;
  p->max_phi_coef = 1.05f;
  p->min_precond = 1.0f;
  p->init_opt_coef = 100.0f;
  p->reference_hpwl = 200000.0f;
  p->min_step_length = 1.0f;
  p->adam_alpha = 100.0f;
  p->adam_beta1 = 0.90f;
  p->adam_beta2 = 0.999f;
  p->adam_epsilon = 1e-8f;
  p->target_density = 1.0f;
  p->scaling_factor = 1.0f;
  p->solver = GP_SOLVER_NESTEROV;
  p->precision = DENS_PREC_BF16;
  p->seed = 1;
  p->log_freq = 50;
}

int gp_destroy(GpHandle* h) {
  if (!h) return GP_OK;
  if (h->s) (void)hipStreamSynchronize(h->s);
  if (h->wl) wl_destroy(h->wl);
  if (h->dens) dens_destroy(h->dens);
  for (void* b : h->bufs) (void)hipFree(b);
  if (h->sc_host) (void)hipHostFree(h->sc_host);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return GP_OK;
}

int gp_create(const LgDesign* d, const GpParams* p, GpHandle** out) {
  if (!d || !p || !out) return fail(GP_ERR_ARG, "null argument");
  *out = nullptr;
  int st = check_params(p);
  if (st) return st;
  if (d->num_insts <= 0 || !d->inst_lx || !d->inst_ly || !d->inst_dx || !d->inst_dy || !d->inst_fixed)
    return fail(GP_ERR_ARG, "no instances");
  if (!(d->core_ux > d->core_lx) || !(d->core_uy > d->core_ly)) return fail(GP_ERR_ARG, "empty core");
  if (d->num_nets < 0 || (d->num_nets > 0 && (!d->net_start || !d->pin_inst || !d->pin_dx || !d->pin_dy)))
    return fail(GP_ERR_ARG, "incomplete netlist");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GP_ERR_DEVICE, "no HIP device");
  const double t0 = now_s();
  GpHandle* h = new GpHandle();
  auto bail = [&](int c) {
    gp_destroy(h);
    return c;
  };
  h->p = *p;
  h->num_insts = d->num_insts;
  h->inst_lx.assign(d->inst_lx, d->inst_lx + d->num_insts);
  h->inst_ly.assign(d->inst_ly, d->inst_ly + d->num_insts);
  // coordinates divided by the scaling factor on import (GPCell(dbInst*,
  // scaling_factor), GPPin, GPDie: SkyPlaceDB.cpp:155-160, GPCell.cpp:45-70)
  const float sf = p->scaling_factor;
  const Die die{(float)d->core_lx / sf, (float)d->core_ly / sf, (float)d->core_ux / sf, (float)d->core_uy / sf};
  h->die = die;
  // importDB (SkyPlaceDB.cpp:140-297)
  std::vector<int> mov_of(d->num_insts, -1);
  std::vector<float> cx, cy, w, hh, npins;
  std::vector<int32_t> macro;
  std::vector<float> frect[4];
  float sum_fixed = 0.f, sum_mov = 0.f, sum_scaled = 0.f;
  const float td = p->target_density;
  for (int i = 0; i < d->num_insts; i++) {
    const float dx0 = (float)d->inst_dx[i], dy0 = (float)d->inst_dy[i];
    const float dx = dx0 / sf, dy = dy0 / sf;
    const float ccx = ((float)d->inst_lx[i] + dx0 / 2) / sf, ccy = ((float)d->inst_ly[i] + dy0 / 2) / sf;
    const bool mac = d->inst_macro && d->inst_macro[i];
    if (d->inst_fixed[i]) {
      const float lx = ccx - dx / 2, ly = ccy - dy / 2, ux = ccx + dx / 2, uy = ccy + dy / 2;
      frect[0].push_back(lx);
      frect[1].push_back(ly);
      frect[2].push_back(ux);
      frect[3].push_back(uy);
      const bool outside = ux <= die.lx || lx >= die.ux || uy <= die.ly || ly >= die.uy;
      if (!outside) sum_fixed += dx * dy;
    } else {
      mov_of[i] = (int)h->mov_inst.size();
      h->mov_inst.push_back(i);
      cx.push_back(ccx);
      cy.push_back(ccy);
      w.push_back(dx);
      hh.push_back(dy);
      macro.push_back(mac ? 1 : 0);
      const float a = dx * dy;
      sum_mov += a;
      sum_scaled += mac ? a * td : a;
    }
  }
  h->n_mov = (int)h->mov_inst.size();
  if (h->n_mov == 0) return bail(fail(GP_ERR_ARG, "no movable instances"));
  const float die_area = (die.ux - die.lx) * (die.uy - die.ly);
  const float util = sum_mov / (die_area - sum_fixed);
  // validateTargetDensity (SkyPlaceDB.cpp:298-317)
  if (td * 100.0 <= util * 100.0)
    return bail(fail(GP_ERR_ARG, "target density (" + std::to_string(td * 100.0) + "%) is not above the utilisation (" +
                                     std::to_string(util * 100.0) + "%)"));
  // createFillers
  std::vector<float> fcx, fcy;
  if ((st = create_fillers(w, hh, die, sum_fixed, sum_scaled, td, h->filler_w, h->filler_h, h->n_fill, fcx, fcy)))
    return bail(st);
  h->n = h->n_mov + h->n_fill;
  const int n = h->n;
  std::vector<int32_t> filler(n, 0);
  for (int k = 0; k < h->n_fill; k++) {
    cx.push_back(fcx[k]);
    cy.push_back(fcy[k]);
    w.push_back(h->filler_w);
    hh.push_back(h->filler_h);
    macro.push_back(0);
    filler[h->n_mov + k] = 1;
  }
  h->w = w;
  h->h = hh;
  // pins (importDB steps 3-5)
  npins.assign(n, 0.f);
  WlNetlist nl{};
  std::vector<int32_t> p2c;
  std::vector<float> ox, oy, px, py;
  const int np = d->num_nets > 0 ? d->net_start[d->num_nets] : 0;
  if (d->num_nets > 0 && d->net_start[0] != 0) return bail(fail(GP_ERR_ARG, "net_start[0] must be 0"));
  p2c.resize(np);
  ox.resize(np);
  oy.resize(np);
  px.resize(np);
  py.resize(np);
  for (int q = 0; q < np; q++) {
    const int i = d->pin_inst[q];
    if (i >= d->num_insts) return bail(fail(GP_ERR_ARG, "pin_inst beyond num_insts"));
    ox[q] = (float)d->pin_dx[q] / sf;
    oy[q] = (float)d->pin_dy[q] / sf;
    if (i < 0) {  // IO pin at its absolute position
      p2c[q] = -1;
      px[q] = ox[q];
      py[q] = oy[q];
      ox[q] = oy[q] = 0.f;
    } else if (mov_of[i] >= 0) {
      const int c = mov_of[i];
      p2c[q] = c;
      npins[c] += 1.f;
      px[q] = cx[c] + ox[q];
      py[q] = cy[c] + oy[q];
    } else {
      p2c[q] = -1;
      const float dx = (float)d->inst_dx[i], dy = (float)d->inst_dy[i];
      px[q] = ((float)d->inst_lx[i] + dx / 2) / sf + ox[q];
      py[q] = ((float)d->inst_ly[i] + dy / 2) / sf + oy[q];
    }
  }
  nl.num_cells = n;
  nl.num_nets = d->num_nets;
  nl.num_pins = np;
  std::vector<int32_t> ns0(1, 0);
  nl.net_start = d->num_nets > 0 ? d->net_start : ns0.data();
  nl.pin2cell = p2c.data();
  nl.pin_offset_x = ox.data();
  nl.pin_offset_y = oy.data();
  nl.pin_x = px.data();
  nl.pin_y = py.data();
  if ((st = wl_status(wl_create(&nl, &h->wl)))) return bail(st);
  // createBins, updateDensitySize, updateFixedOverlapArea
  DensDesign dd{};
  dd.die_lx = die.lx;
  dd.die_ly = die.ly;
  dd.die_ux = die.ux;
  dd.die_uy = die.uy;
  dd.target_density = td;
  dd.num_movable = n;
  dd.cell_dx = w.data();
  dd.cell_dy = hh.data();
  dd.is_macro = macro.data();
  dd.is_filler = filler.data();
  dd.num_fixed = (int)frect[0].size();
  dd.fixed_lx = frect[0].data();
  dd.fixed_ly = frect[1].data();
  dd.fixed_ux = frect[2].data();
  dd.fixed_uy = frect[3].data();
  dd.num_bin_x = p->num_bin_x;
  dd.num_bin_y = p->num_bin_y;
  int32_t shape[2] = {0, 0};
  if ((st = dens_status(dens_prepare(&dd, shape, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     nullptr, nullptr, nullptr))))
    return bail(st);
  h->nbx = shape[0];
  h->nby = shape[1];
  const size_t B = (size_t)h->nbx * h->nby;
  std::vector<float> blx(h->nbx), bux(h->nbx), bly(h->nby), buy(h->nby), fa(B), sa(B), cdw(n), cdh(n), csc(n), scal(3);
  if ((st = dens_status(dens_prepare(&dd, shape, blx.data(), bux.data(), bly.data(), buy.data(), fa.data(), sa.data(),
                                     cdw.data(), cdh.data(), csc.data(), scal.data()))))
    return bail(st);
  DensGeom g{};
  g.num_bin_x = h->nbx;
  g.num_bin_y = h->nby;
  g.die_lx = die.lx;
  g.die_ly = die.ly;
  g.die_ux = die.ux;
  g.die_uy = die.uy;
  g.bin_w = scal[0];
  g.bin_h = scal[1];
  g.target_density = td;
  g.bin_lx = blx.data();
  g.bin_ux = bux.data();
  g.bin_ly = bly.data();
  g.bin_uy = buy.data();
  g.fixed_area = fa.data();
  g.scaled_bin_area = sa.data();
  g.num_movable = n;
  g.cell_dw = cdw.data();
  g.cell_dh = cdh.data();
  g.cell_scale = csc.data();
  g.is_filler = filler.data();
  g.is_macro = macro.data();
  g.sum_movable_area = scal[2];
  g.precision = p->precision;
  if ((st = dens_status(dens_create(&g, &h->dens)))) return bail(st);
  // TargetFunction::TargetFunction (TargetFunction.cu:17-45)
  h->lambda = p->init_lambda;
  h->gamma_inv = 2.0f * p->init_gamma_inv_coef / (scal[0] + scal[1]);
  // device state
  GP_TRY(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  const size_t V = 2 * (size_t)n;
  float** vecs[] = {&h->cur_var, &h->new_var, &h->cur_pre, &h->new_pre, &h->prv_pre, &h->cur_grad, &h->new_grad,
                    &h->prv_grad, &h->wl_grad, &h->dens_grad};
  for (float** v : vecs)
    if ((st = dalloc(h, v, V))) return bail(st);
  if (p->solver == GP_SOLVER_ADAM) {
    for (float** v : {&h->m1, &h->v1, &h->m2, &h->v2})
      if ((st = dalloc(h, v, V))) return bail(st);
  }
  if ((st = dupload(h, &h->dw, w.data(), n)) || (st = dupload(h, &h->dh, hh.data(), n)) ||
      (st = dupload(h, &h->num_pins, npins.data(), n)) || (st = dalloc(h, &h->part, 2 * (size_t)kRedBlocks)) ||
      (st = dalloc(h, &h->hpwl_acc, 1)) || (st = dalloc(h, &h->sc_dev, 1)))
    return bail(st);
  if (hipHostMalloc((void**)&h->sc_host, sizeof(Scalars)) != hipSuccess)
    return bail(fail(GP_ERR_MEMORY, "hipHostMalloc"));
  std::vector<float> pos(V);
  std::copy(cx.begin(), cx.end(), pos.begin());
  std::copy(cy.begin(), cy.end(), pos.begin() + n);
  GP_TRY(hipMemcpy(h->cur_var, pos.data(), V * sizeof(float), hipMemcpyHostToDevice));
  GP_TRY(hipMemcpy(h->cur_pre, pos.data(), V * sizeof(float), hipMemcpyHostToDevice));
  // HPWL of the design's own positions
  if ((st = wl_status(wl_compute_hpwl(h->wl, &h->hpwl_initial, h->s)))) return bail(st);
  if (p->reference_hpwl_ratio > 0.f) h->p.reference_hpwl = (float)(p->reference_hpwl_ratio * h->hpwl_initial);
  if (!(h->p.reference_hpwl > 0.f)) return bail(fail(GP_ERR_ARG, "reference HPWL must be positive"));
  h->t_init = now_s() - t0;
  *out = h;
  return GP_OK;
}

int gp_initial_place(GpHandle* h) {
  if (!h) return fail(GP_ERR_ARG, "null handle");
  const double t0 = now_s();
  const Die& d = h->die;
  const double mx = 0.5 * ((double)d.lx + (double)d.ux), my = 0.5 * ((double)d.ly + (double)d.uy);
  const double sx = ((double)d.ux - (double)d.lx) * 0.05, sy = ((double)d.uy - (double)d.ly) * 0.05;
  hipLaunchKernelGGL(k_init_place, dim3(blocks(h->n)), dim3(kThreads), 0, h->s, h->n, h->n_mov, h->p.seed, mx, my,
                     sx, sy, d, h->dw, h->dh, h->cur_var);
  GP_TRY(hipGetLastError());
  GP_TRY(hipMemcpyAsync(h->cur_pre, h->cur_var, 2 * (size_t)h->n * sizeof(float), hipMemcpyDeviceToDevice, h->s));
  GP_TRY(hipStreamSynchronize(h->s));
  h->solver_ready = false;
  h->t_init += now_s() - t0;
  return GP_OK;
}

int gp_init_solver(GpHandle* h) {
  if (!h) return fail(GP_ERR_ARG, "null handle");
  const double t0 = now_s();
  int st;
  h->iter = 0;
  h->done = h->converged = false;
  h->evals = 0;
  h->lambda = h->p.init_lambda;
  if (h->p.solver == GP_SOLVER_NESTEROV) {
    // NesterovSolver::initSolver (NesterovSolver.cu:63-90)
    GP_TRY(hipMemcpyAsync(h->cur_pre, h->cur_var, 2 * (size_t)h->n * sizeof(float), hipMemcpyDeviceToDevice, h->s));
    if ((st = initial_grad(h, h->cur_pre, h->cur_grad))) return st;
    hipLaunchKernelGGL(k_move_backward, dim3(blocks(2LL * h->n)), dim3(kThreads), 0, h->s, h->n, h->p.init_opt_coef,
                       h->cur_pre, h->cur_grad, h->prv_pre, h->dw, h->dh, h->die);
    if ((st = evaluate(h, h->prv_pre, h->prv_grad, h->lambda, h->cur_pre, h->cur_grad))) return st;
    h->step = lipschitz(*h->sc_host, 2 * h->n);
    update_parameters(h);
    h->cur_a = 1.0f;
  } else {
    // AdamSolver::initSolver (AdamSolver.cu:50-65)
    h->alpha = h->p.adam_alpha;
    h->beta1k = h->p.adam_beta1;
    h->beta2k = h->p.adam_beta2;
    if ((st = initial_grad(h, h->cur_var, h->cur_grad))) return st;
    // HPWL at the initial point for updateParameters
    if ((st = wl_status(wl_compute_hpwl_async(h->wl, h->hpwl_acc, h->s)))) return st;
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, h->s, h->part, 0, h->hpwl_acc,
                       dens_get_overflow_device(h->dens), h->sc_dev);
    GP_TRY(hipMemcpyAsync(h->sc_host, h->sc_dev, sizeof(Scalars), hipMemcpyDeviceToHost, h->s));
    GP_TRY(hipStreamSynchronize(h->s));
    update_parameters(h);
    GP_TRY(hipMemsetAsync(h->m1, 0, 2 * (size_t)h->n * sizeof(float), h->s));
    GP_TRY(hipMemsetAsync(h->v1, 0, 2 * (size_t)h->n * sizeof(float), h->s));
  }
  GP_TRY(hipStreamSynchronize(h->s));
  h->solver_ready = true;
  if (h->p.verbose) {
    std::printf("Initial HPWL     : %-6.3f\n", h->hpwl / 1e6);
    std::printf("Initial Overflow : %-3.3f\n", h->overflow);
    std::printf("\n----------------------------------------------\n");
    std::printf("|  Iter |     HPWL     | Overflow | Time (s) |\n");
    std::printf("----------------------------------------------\n");
  }
  h->t_solve += now_s() - t0;
  return GP_OK;
}

int gp_iterate(GpHandle* h, int32_t count, int32_t* done) {
  if (!h) return fail(GP_ERR_ARG, "null handle");
  if (!h->solver_ready) return fail(GP_ERR_STATE, "gp_init_solver first");
  const double t0 = now_s();
  for (int k = 0; k < count && !h->done; k++) {
    if (h->iter >= h->p.max_opt_iter) {
      h->done = true;
      break;
    }
    int st = h->p.solver == GP_SOLVER_NESTEROV ? nesterov_iteration(h) : adam_iteration(h);
    if (st) return st;
    if (check_convergence(h)) {
      h->converged = h->done = true;
      break;
    }
    // iterEndCbk (TargetFunction.cu:330-343)
    update_parameters(h);
    record(h, h->iter, h->last_bt);
    if (h->iter == 0 || (h->p.log_freq > 0 && h->iter % h->p.log_freq == 0))
      progress(h, h->iter, h->t_solve + now_s() - t0);
    h->iter++;
  }
  if (h->iter >= h->p.max_opt_iter) h->done = true;
  h->t_solve += now_s() - t0;
  if (done) *done = h->done ? 1 : 0;
  return GP_OK;
}

int gp_solve(GpHandle* h, GpTrace* trace) {
  if (!h) return fail(GP_ERR_ARG, "null handle");
  h->trace = trace ? *trace : GpTrace{};
  int st = gp_init_solver(h);
  if (st) return st;
  if (h->p.verbose) std::printf("%s Solve Start\n", h->p.solver == GP_SOLVER_NESTEROV ? "Nesterov" : "Adam");
  int32_t done = 0;
  while (!done)
    if ((st = gp_iterate(h, 1 << 30, &done))) return st;
  if (h->p.verbose) {
    progress(h, h->iter, h->t_solve);
    std::printf("----------------------------------------------\n");
    std::printf(" HPWL     | %-10.5f\n", h->hpwl);
    std::printf(" Iter     | %-10d  \n", h->iter);
    std::printf(" Overflow | %-10.5f\n", h->overflow);
    std::printf("----------------------------------------------\n");
  }
  return GP_OK;
}

int gp_get_positions(GpHandle* h, float* cx, float* cy) {
  if (!h) return fail(GP_ERR_ARG, "null handle");
  GP_TRY(hipStreamSynchronize(h->s));
  if (cx) GP_TRY(hipMemcpy(cx, h->cur_var, (size_t)h->n * sizeof(float), hipMemcpyDeviceToHost));
  if (cy) GP_TRY(hipMemcpy(cy, h->cur_var + h->n, (size_t)h->n * sizeof(float), hipMemcpyDeviceToHost));
  return GP_OK;
}

int gp_set_positions(GpHandle* h, const float* cx, const float* cy) {
  if (!h || !cx || !cy) return fail(GP_ERR_ARG, "null argument");
  GP_TRY(hipStreamSynchronize(h->s));
  const size_t b = (size_t)h->n * sizeof(float);
  GP_TRY(hipMemcpy(h->cur_var, cx, b, hipMemcpyHostToDevice));
  GP_TRY(hipMemcpy(h->cur_var + h->n, cy, b, hipMemcpyHostToDevice));
  GP_TRY(hipMemcpy(h->cur_pre, h->cur_var, 2 * b, hipMemcpyDeviceToDevice));
  h->solver_ready = false;
  return GP_OK;
}

int gp_get_state(GpHandle* h, float* scalars, float* cur_var, float* cur_pre, float* cur_grad) {
  if (!h) return fail(GP_ERR_ARG, "null handle");
  GP_TRY(hipStreamSynchronize(h->s));
  if (scalars) {
    const float v[8] = {h->lambda, h->gamma_inv, h->step, h->cur_a, h->hpwl, h->prev_hpwl, h->overflow,
                        (float)h->iter};
    std::memcpy(scalars, v, sizeof(v));
  }
  const size_t b = 2 * (size_t)h->n * sizeof(float);
  if (cur_var) GP_TRY(hipMemcpy(cur_var, h->cur_var, b, hipMemcpyDeviceToHost));
  if (cur_pre) GP_TRY(hipMemcpy(cur_pre, h->cur_pre, b, hipMemcpyDeviceToHost));
  if (cur_grad) GP_TRY(hipMemcpy(cur_grad, h->cur_grad, b, hipMemcpyDeviceToHost));
  return GP_OK;
}

int gp_export(GpHandle* h, int32_t* out_lx, int32_t* out_ly) {
  if (!h || !out_lx || !out_ly) return fail(GP_ERR_ARG, "null argument");
  std::vector<float> cx(h->n), cy(h->n);
  int st = gp_get_positions(h, cx.data(), cy.data());
  if (st) return st;
  std::copy(h->inst_lx.begin(), h->inst_lx.end(), out_lx);
  std::copy(h->inst_ly.begin(), h->inst_ly.end(), out_ly);
  for (int c = 0; c < h->n_mov; c++) {
    const int i = h->mov_inst[c];
    const float lx = cx[c] - h->w[c] / 2, ly = cy[c] - h->h[c] / 2;  // GPCell::lx() / ly()
    out_lx[i] = (int32_t)std::round(lx * h->p.scaling_factor);
    out_ly[i] = (int32_t)std::round(ly * h->p.scaling_factor);
  }
  return GP_OK;
}

int gp_get_stats(GpHandle* h, GpStats* s) {
  if (!h || !s) return fail(GP_ERR_ARG, "null argument");
  GpStats r{};
  r.num_movable = h->n_mov;
  r.num_fillers = h->n_fill;
  r.num_bin_x = h->nbx;
  r.num_bin_y = h->nby;
  r.filler_width = h->filler_w;
  r.filler_height = h->filler_h;
  r.iterations = h->iter;
  r.converged = h->converged ? 1 : 0;
  r.evaluations = h->evals;
  r.hpwl = h->hpwl;
  r.overflow = h->overflow;
  r.lambda = h->lambda;
  r.gamma_inv = h->gamma_inv;
  r.reference_hpwl = h->p.reference_hpwl;
  r.hpwl_initial_db = h->hpwl_initial;
  // SkyPlaceDB::updateHpwl at the final centres
  int st = wl_status(wl_update_pin_coordinates(h->wl, h->cur_var, h->cur_var + h->n, h->s));
  if (st) return st;
  if ((st = wl_status(wl_compute_hpwl(h->wl, &r.hpwl_db, h->s)))) return st;
  r.t_init_s = h->t_init;
  r.t_solve_s = h->t_solve;
  *s = r;
  return GP_OK;
}

int gp_place_design(const LgDesign* d, const GpParams* p, int32_t* out_lx, int32_t* out_ly, GpStats* stats,
                    GpTrace* trace) {
  if (!out_lx || !out_ly) return fail(GP_ERR_ARG, "null output");
  GpHandle* h = nullptr;
  int st = gp_create(d, p, &h);
  if (st) return st;
  if (p->verbose)
    std::printf("global_place: %d movable, %d fillers (%.1f x %.1f), bins %d x %d\n", h->n_mov, h->n_fill,
                h->filler_w, h->filler_h, h->nbx, h->nby);
  if ((st = gp_initial_place(h))) return gp_destroy(h), st;
  if (!p->initial_place_only && (st = gp_solve(h, trace))) return gp_destroy(h), st;
  if ((st = gp_export(h, out_lx, out_ly))) return gp_destroy(h), st;
  if (stats && (st = gp_get_stats(h, stats))) return gp_destroy(h), st;
  gp_destroy(h);
  return GP_OK;
}

}  // extern "C"
