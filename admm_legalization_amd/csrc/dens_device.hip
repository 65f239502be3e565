// Electrostatic density path on MI355X: tiled LDS scatter, overflow, Poisson
// solve as cosine / sine GEMMs on MFMA, gather.  C ABI in skyline_density.h.
//
// Data layout in HBM (bins y-major, bin = x + y * nbx):
//   cgeo[N]   float4 {density width, density height, scatter scale, gather scale}
//   cfill[N]  int    filler flag
//   slab[T][2][SW*SW]  per-tile movable / filler partial areas (SW = 32 + halo)
//   big[2][B] uint64   fixed-point sums of the cells wider than the halo
//   mov, fil, dens, psi, ex, ey, lambda, fixed, scaled  [B] float
//
// Scatter (computeDensityContributionOfEachCell, DensityGradient.cu:61-138):
// cells are bucketed by the 32x32-bin tile holding their lower-left footprint
// bin; one workgroup per tile accumulates its cells in LDS (tile plus a halo
// of H bins up and right), then writes the tile's slab with plain stores; a
// per-bin resolve pass adds a bin's own slab entry and the halo entries of the
// (up to three) tiles below / left of it, and computes density and overflow in
// the same pass.  Cells wider than the halo go to a separate list whose bins
// are added with global atomics into a per-bin buffer that the resolve pass
// adds and clears.  All sums are 64-bit fixed point (overlap areas scaled by a
// power of two chosen from the bin area), and the overflow is reduced in
// block order: an evaluation's results do not depend on the order in which
// the hardware runs the atomics, so equal inputs give equal bits.
//
// Poisson (PoissonSolver::solvePoisson, PoissonSolver.cu:224-302): with
// M = nby rows, N = nbx columns and k-contiguous operands (out = A * B^T):
//   U[N x M]   = CN   * rho^T                                (CN[j,n] = cos(pi j (2n+1)/2N))
//   A'[M x N]  = 4/(MN) CM * U^T, divided by wy^2 + wx^2,  A'[0,0] = 0
//   [V;Vx][2N x M] = [DNt; SNw] * A'^T
//   [psi;Ey][2M x N] = [DMt; SMw] * V^T,   Ex[M x N] = DMt * Vx^T
// with DNt[n,w] = c_w cos(pi w (2n+1)/2N) (c_0 = 1, else 2), SNw[n,w] =
// 2 sin(pi w (2n+1)/2N) * 2 pi w / N (w >= 1), and the same along y.  These are
// the reference's DCT-II (scaled 4/(MN)), DCT-III, and IDXST transforms with
// multiplyW folded into the sine matrices.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/skyline_density.h"

namespace densdev {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define DENS_TRY(x)                                                              \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      return fail(DENS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTile = 32;   // bins per tile side
constexpr int kResolveBlocks = 4096;
constexpr int kChunk = 2048;  // records per accumulation workgroup
constexpr int kMaxHalo = 8;
constexpr int kBK = 64;     // GEMM k-step

// ---------------------------------------------------------------------------
// GEMM: C = alpha * A * B^T, A [R x K], B [Cn x K], both k-contiguous.  Rows
// below `split` go to C0, the others to C1.  Optional epilogue divides by
// wy2[row] + wx2[col] (zero at (0,0)).  256 threads = 2 x 2 waves, each wave
// owns a (BM/2) x (BN/2) block of 32x32 MFMA accumulators.
// ---------------------------------------------------------------------------
struct Gemm {
  const float* A;
  const float* B;
  float* C0;
  float* C1;
  int lda, ldb, ldc, split;
  int R, Cn, K;
  float alpha;
  const float* wy2;
  const float* wx2;
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int PREC, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_nt(Gemm g) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 32, FN = WN / 32;
  constexpr int LD16 = kBK + 8;  // bf16 row stride (80 B: conflict-free 16-B reads)
  constexpr int LDF = kBK + 4;   // f32 row stride
  constexpr int AF4 = BM * kBK / 4 / 256;
  constexpr int BF4 = BN * kBK / 4 / 256;
  constexpr bool F32 = PREC == DENS_PREC_F32;
  constexpr bool SPLIT = PREC == DENS_PREC_SPLIT_BF16;

  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* sAh = reinterpret_cast<__bf16*>(smem);
  __bf16* sAl = sAh + BM * LD16;
  __bf16* sBh = sAl + BM * LD16;
  __bf16* sBl = sBh + BN * LD16;
  float* sAf = reinterpret_cast<float*>(smem);
  float* sBf = sAf + BM * LDF;

  const int tiles_n = (g.Cn + BN - 1) / BN;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x - tm * tiles_n;
  const int r0 = tm * BM, c0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;

  float4 ra[AF4], rb[BF4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < AF4; i++) {
      int f = tid + i * 256, row = f / (kBK / 4), kq = f - row * (kBK / 4);
      int gr = r0 + row, gk = k0 + 4 * kq;
      ra[i] = (gr < g.R && gk < g.K) ? *reinterpret_cast<const float4*>(g.A + (size_t)gr * g.lda + gk)
                                     : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < BF4; i++) {
      int f = tid + i * 256, row = f / (kBK / 4), kq = f - row * (kBK / 4);
      int gr = c0 + row, gk = k0 + 4 * kq;
      rb[i] = (gr < g.Cn && gk < g.K) ? *reinterpret_cast<const float4*>(g.B + (size_t)gr * g.ldb + gk)
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  auto put16 = [&](__bf16* hi, __bf16* lo, int row, int kq, float4 v) {
    bf16x4 h, l;
    __bf16 a, b;
    split_bf16(v.x, a, b); h[0] = a; l[0] = b;
    split_bf16(v.y, a, b); h[1] = a; l[1] = b;
    split_bf16(v.z, a, b); h[2] = a; l[2] = b;
    split_bf16(v.w, a, b); h[3] = a; l[3] = b;
    *reinterpret_cast<bf16x4*>(hi + row * LD16 + 4 * kq) = h;
    if (SPLIT) *reinterpret_cast<bf16x4*>(lo + row * LD16 + 4 * kq) = l;
  };
  auto sstore = [&]() {
#pragma unroll
    for (int i = 0; i < AF4; i++) {
      int f = tid + i * 256, row = f / (kBK / 4), kq = f - row * (kBK / 4);
      if (F32) *reinterpret_cast<float4*>(sAf + row * LDF + 4 * kq) = ra[i];
      else put16(sAh, sAl, row, kq, ra[i]);
    }
#pragma unroll
    for (int i = 0; i < BF4; i++) {
      int f = tid + i * 256, row = f / (kBK / 4), kq = f - row * (kBK / 4);
      if (F32) *reinterpret_cast<float4*>(sBf + row * LDF + 4 * kq) = rb[i];
      else put16(sBh, sBl, row, kq, rb[i]);
    }
  };

  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

  const int l32 = lane & 31, lhi = lane >> 5;
  gload(0);
  for (int k0 = 0; k0 < g.K; k0 += kBK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + kBK < g.K) gload(k0 + kBK);
    if (F32) {
#pragma unroll
      for (int kk = 0; kk < kBK; kk += 2) {
        float a[FM], b[FN];
#pragma unroll
        for (int i = 0; i < FM; i++) a[i] = sAf[(wr * WM + i * 32 + l32) * LDF + kk + lhi];
#pragma unroll
        for (int j = 0; j < FN; j++) b[j] = sBf[(wc * WN + j * 32 + l32) * LDF + kk + lhi];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < kBK; kk += 16) {
        bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; i++) {
          int off = (wr * WM + i * 32 + l32) * LD16 + kk + 8 * lhi;
          ah[i] = *reinterpret_cast<const bf16x8*>(sAh + off);
          if (SPLIT) al[i] = *reinterpret_cast<const bf16x8*>(sAl + off);
        }
#pragma unroll
        for (int j = 0; j < FN; j++) {
          int off = (wc * WN + j * 32 + l32) * LD16 + kk + 8 * lhi;
          bh[j] = *reinterpret_cast<const bf16x8*>(sBh + off);
          if (SPLIT) bl[j] = *reinterpret_cast<const bf16x8*>(sBl + off);
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) {
            if (SPLIT) {
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            }
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
      }
    }
  }

  // accumulator layout of a 32x32 block: lane l, register r ->
  // row 8*(r/4) + 4*(l/32) + r%4, column l%32
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        int row = r0 + wr * WM + i * 32 + 8 * (r >> 2) + 4 * lhi + (r & 3);
        int col = c0 + wc * WN + j * 32 + l32;
        if (row < g.R && col < g.Cn) {
          float v = g.alpha * acc[i][j][r];
          if (g.wy2) v = (row == 0 && col == 0) ? 0.f : v / (g.wy2[row] + g.wx2[col]);
          float* dst = row < g.split ? g.C0 + (size_t)row * g.ldc : g.C1 + (size_t)(row - g.split) * g.ldc;
          dst[col] = v;
        }
      }
}

template <int PREC, int BM, int BN>
constexpr size_t gemm_lds() {
  return PREC == DENS_PREC_F32 ? (size_t)(BM + BN) * (kBK + 4) * 4 : (size_t)(BM + BN) * (kBK + 8) * 2 * 2;
}

// Tile shapes of the GEMM kernels (BM x BN): 1 = 64x64, 2 = 64x128,
// 3 = 128x64, 4 = 128x128, 5 = 256x128; 0 = by problem size (128x128 from
// 512 tiles).  Every shape accumulates each output element over k in the
// same order, so the shape changes the speed only.
constexpr int kNumTiles = 5;

template <int PREC, int BM, int BN>
void launch_gemm_tile(const Gemm& g, hipStream_t s) {
  const int grid = ((g.R + BM - 1) / BM) * ((g.Cn + BN - 1) / BN);
  constexpr size_t lds = gemm_lds<PREC, BM, BN>();
  hipLaunchKernelGGL((k_gemm_nt<PREC, BM, BN>), dim3(grid), dim3(256), lds, s, g);
}

template <int PREC>
int launch_gemm_p(const Gemm& g, hipStream_t s, int tile) {
  if (tile == 0) {
    const long tiles128 = (long)((g.R + 127) / 128) * ((g.Cn + 127) / 128);
    tile = tiles128 >= 512 ? 4 : 1;
  }
  switch (tile) {
    case 1: launch_gemm_tile<PREC, 64, 64>(g, s); break;
    case 2: launch_gemm_tile<PREC, 64, 128>(g, s); break;
    case 3: launch_gemm_tile<PREC, 128, 64>(g, s); break;
    case 4: launch_gemm_tile<PREC, 128, 128>(g, s); break;
    default: launch_gemm_tile<PREC, 256, 128>(g, s); break;
  }
  DENS_TRY(hipGetLastError());
  return DENS_OK;
}

// ---------------------------------------------------------------------------
// GEMM on operands already in bf16 (BF16 and split-BF16 chains): C = alpha *
// A * B^T with A = Ah (+ Al), B = Bh (+ Bl), k-contiguous, K a multiple of 8.
// The constant cosine / sine matrices are split once at poisson_create and
// every intermediate of the chain is split by the epilogue of the GEMM that
// produces it, so the k-loop only moves bf16: 16-byte global loads into
// registers for the next k-step while the MFMAs run on the current one, two
// LDS stages, one barrier per k-step.  Per 16-deep k chunk each 32 x 32
// block takes al*bh, ah*bl, ah*bh (split) or ah*bh, in that order, as the
// FP32-operand kernel does, and the epilogue splits the same FP32 values,
// so both kernels give the same bits.  256 threads = 2 x 2 waves, each
// (BM/2) x (BN/2) of 32 x 32 MFMA blocks.
// ---------------------------------------------------------------------------
struct GemmBf {
  const __bf16* Ah;
  const __bf16* Al;
  const __bf16* Bh;
  const __bf16* Bl;
  float* C0;          // FP32 output rows [0, split) ...
  float* C1;          // ... and [split, R), or
  __bf16* Ch;         // bf16 hi / lo output (the next GEMM's operand), all rows
  __bf16* Cl;
  int lda, ldb, ldc, split;
  int R, Cn, K;
  float alpha;
  const float* wy2;
  const float* wx2;
};

// k-step BK: 64, or 32 for the split tiles of 128 x 128 and up, whose two
// LDS stages at 64 would leave one workgroup per CU
template <bool SPLIT, int BM, int BN>
constexpr int bf_bk() {
  return (SPLIT && BM * BN >= 128 * 128) ? 32 : 64;
}

template <bool SPLIT, int BM, int BN>
__device__ __forceinline__ void gemm_bf_body(const GemmBf& g, int bid) {
  constexpr int BK = bf_bk<SPLIT, BM, BN>();
  constexpr int CPR = BK / 8;              // 16-byte chunks per operand row of a k-step
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 32, FN = WN / 32;
  constexpr int LD = BK + 8;               // bf16 row stride: 16 B of padding, conflict-free 16-B reads
  constexpr int CA = BM * BK / 8 / 256;    // 16-byte chunks per thread, A (per part)
  constexpr int CB = BN * BK / 8 / 256;
  constexpr int NP = SPLIT ? 2 : 1;        // operand parts: hi (+ lo)
  constexpr int STAGE = (BM + BN) * LD * NP;  // bf16 elements per LDS stage
  // k-steps whose operands are in flight in registers: the small tiles (one
  // workgroup per CU at 1024^2, little MFMA work per step) load 3 steps
  // ahead, the large ones 1 or 2
  constexpr int PD = BM * BN <= 64 * 128 ? 3 : (BM * BN <= 128 * 128 ? 2 : 1);
  static_assert(CA >= 1 && CB >= 1 && CA * 256 * 8 == BM * BK && CB * 256 * 8 == BN * BK,
                "tile does not split into 16-byte chunks over 256 threads");

  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* lds = reinterpret_cast<__bf16*>(smem);

  const int tiles_n = (g.Cn + BN - 1) / BN;
  const int tm = bid / tiles_n, tn = bid - tm * tiles_n;
  const int r0 = tm * BM, c0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int l32 = lane & 31, lhi = lane >> 5;

  typedef int int4v __attribute__((ext_vector_type(4)));
  // PD register sets; set u holds the k-steps congruent to u mod PD (every
  // index below is a constant once the step loop is unrolled)
  int4v ra[PD][NP][CA], rb[PD][NP][CB];
  auto gload = [&](int4v (&xa)[NP][CA], int4v (&xb)[NP][CB], int k0) {
#pragma unroll
    for (int i = 0; i < CA; i++) {
      const int f = tid + i * 256, row = f / CPR, kc = (f % CPR) * 8;
      const int gr = r0 + row, gk = k0 + kc;
      const bool ok = gr < g.R && gk < g.K;
      const size_t o = ok ? (size_t)gr * g.lda + gk : 0;
      xa[0][i] = ok ? *reinterpret_cast<const int4v*>(g.Ah + o) : int4v{0, 0, 0, 0};
      if (SPLIT) xa[NP - 1][i] = ok ? *reinterpret_cast<const int4v*>(g.Al + o) : int4v{0, 0, 0, 0};
    }
#pragma unroll
    for (int i = 0; i < CB; i++) {
      const int f = tid + i * 256, row = f / CPR, kc = (f % CPR) * 8;
      const int gr = c0 + row, gk = k0 + kc;
      const bool ok = gr < g.Cn && gk < g.K;
      const size_t o = ok ? (size_t)gr * g.ldb + gk : 0;
      xb[0][i] = ok ? *reinterpret_cast<const int4v*>(g.Bh + o) : int4v{0, 0, 0, 0};
      if (SPLIT) xb[NP - 1][i] = ok ? *reinterpret_cast<const int4v*>(g.Bl + o) : int4v{0, 0, 0, 0};
    }
  };
  // stage layout: A hi [BM][LD], B hi [BN][LD], then A lo, B lo
  auto sstore = [&](int st, const int4v (&xa)[NP][CA], const int4v (&xb)[NP][CB]) {
    __bf16* base = lds + st * STAGE;
#pragma unroll
    for (int p = 0; p < NP; p++) {
      __bf16* sa = base + p * (BM + BN) * LD;
      __bf16* sb = sa + BM * LD;
#pragma unroll
      for (int i = 0; i < CA; i++) {
        const int f = tid + i * 256, row = f / CPR, kc = (f % CPR) * 8;
        *reinterpret_cast<int4v*>(sa + row * LD + kc) = xa[p][i];
      }
#pragma unroll
      for (int i = 0; i < CB; i++) {
        const int f = tid + i * 256, row = f / CPR, kc = (f % CPR) * 8;
        *reinterpret_cast<int4v*>(sb + row * LD + kc) = xb[p][i];
      }
    }
  };

  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

  const int nk = (g.K + BK - 1) / BK;
#pragma unroll
  for (int u = 0; u < PD; u++)
    if (u < nk) gload(ra[u], rb[u], u * BK);
  sstore(0, ra[0], rb[0]);
  __syncthreads();
  for (int ks0 = 0; ks0 < nk; ks0 += PD)
#pragma unroll
  for (int u = 0; u < PD; u++) {
    const int ks = ks0 + u;
    if (ks >= nk) break;
    // set u's step ks is in LDS: the set takes step ks + PD
    if (ks + PD < nk) gload(ra[u], rb[u], (ks + PD) * BK);
    const bool more = ks + 1 < nk;
    const __bf16* base = lds + (ks & 1) * STAGE;
    const __bf16* sah = base;
    const __bf16* sbh = base + BM * LD;
    const __bf16* sal = base + (BM + BN) * LD;
    const __bf16* sbl = sal + BM * LD;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) {
        const int off = (wr * WM + i * 32 + l32) * LD + kk + 8 * lhi;
        ah[i] = *reinterpret_cast<const bf16x8*>(sah + off);
        if (SPLIT) al[i] = *reinterpret_cast<const bf16x8*>(sal + off);
      }
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int off = (wc * WN + j * 32 + l32) * LD + kk + 8 * lhi;
        bh[j] = *reinterpret_cast<const bf16x8*>(sbh + off);
        if (SPLIT) bl[j] = *reinterpret_cast<const bf16x8*>(sbl + off);
      }
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) {
          if (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    // the other stage was last read in step ks - 1, which every wave left
    // at that step's barrier; step ks + 1 is in set (u + 1) mod PD
    if (more) sstore((ks + 1) & 1, ra[(u + 1) % PD], rb[(u + 1) % PD]);
    __syncthreads();
  }

  // accumulator of a 32 x 32 block: lane l, register r -> row 8*(r/4) +
  // 4*(l/32) + r%4, column l%32
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row = r0 + wr * WM + i * 32 + 8 * (r >> 2) + 4 * lhi + (r & 3);
        const int col = c0 + wc * WN + j * 32 + l32;
        if (row < g.R && col < g.Cn) {
          float v = g.alpha * acc[i][j][r];
          if (g.wy2) v = (row == 0 && col == 0) ? 0.f : v / (g.wy2[row] + g.wx2[col]);
          if (g.Ch) {
            __bf16 hi, lo;
            split_bf16(v, hi, lo);
            g.Ch[(size_t)row * g.ldc + col] = hi;
            if (SPLIT) g.Cl[(size_t)row * g.ldc + col] = lo;
          } else {
            float* dst = row < g.split ? g.C0 + (size_t)row * g.ldc : g.C1 + (size_t)(row - g.split) * g.ldc;
            dst[col] = v;
          }
        }
      }
}

template <bool SPLIT, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_bf(GemmBf g) {
  gemm_bf_body<SPLIT, BM, BN>(g, blockIdx.x);
}

// Two independent GEMMs of one tile shape in one launch: workgroups
// [0, nb0) compute g0's tiles, the rest g1's.
template <bool SPLIT, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_bf2(GemmBf g0, GemmBf g1, int nb0) {
  if ((int)blockIdx.x < nb0) gemm_bf_body<SPLIT, BM, BN>(g0, blockIdx.x);
  else gemm_bf_body<SPLIT, BM, BN>(g1, blockIdx.x - nb0);
}

// FP32 -> bf16 hi (+ lo) of n values (the chain's input density)
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_split_bf16(const float* __restrict__ x, __bf16* hi, __bf16* lo, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    __bf16 h, l;
    split_bf16(x[i], h, l);
    hi[i] = h;
    if (SPLIT) lo[i] = l;
  }
}

template <bool SPLIT, int BM, int BN>
constexpr size_t gemm_bf_lds() {
  return (size_t)2 * (BM + BN) * (bf_bk<SPLIT, BM, BN>() + 8) * (SPLIT ? 2 : 1) * 2;
}

template <bool SPLIT, int BM, int BN>
void launch_gemm_bf_tile(const GemmBf& g, const GemmBf* g1, hipStream_t s) {
  const int grid = ((g.R + BM - 1) / BM) * ((g.Cn + BN - 1) / BN);
  constexpr size_t lds = gemm_bf_lds<SPLIT, BM, BN>();
  if (g1) {
    const int grid1 = ((g1->R + BM - 1) / BM) * ((g1->Cn + BN - 1) / BN);
    hipLaunchKernelGGL((k_gemm_bf2<SPLIT, BM, BN>), dim3(grid + grid1), dim3(256), lds, s, g, *g1, grid);
  } else {
    hipLaunchKernelGGL((k_gemm_bf<SPLIT, BM, BN>), dim3(grid), dim3(256), lds, s, g);
  }
}

template <bool SPLIT>
int launch_gemm_bf_p(const GemmBf& g, const GemmBf* g1, hipStream_t s, int tile) {
  switch (tile) {
    case 1: launch_gemm_bf_tile<SPLIT, 64, 64>(g, g1, s); break;
    case 2: launch_gemm_bf_tile<SPLIT, 64, 128>(g, g1, s); break;
    case 3: launch_gemm_bf_tile<SPLIT, 128, 64>(g, g1, s); break;
    case 4: launch_gemm_bf_tile<SPLIT, 128, 128>(g, g1, s); break;
    default: launch_gemm_bf_tile<SPLIT, 256, 128>(g, g1, s); break;
  }
  DENS_TRY(hipGetLastError());
  return DENS_OK;
}

// g1 (optional): a second GEMM, independent of g, computed in the same launch
int launch_gemm_bf(int prec, const GemmBf& g, hipStream_t s, int tile, const GemmBf* g1 = nullptr) {
  if (g.K % 8 || g.lda % 8 || g.ldb % 8) return fail(DENS_ERR_ARG, "bf16 GEMM operands must be multiples of 8 wide");
  if (g1 && (g1->K % 8 || g1->lda % 8 || g1->ldb % 8))
    return fail(DENS_ERR_ARG, "bf16 GEMM operands must be multiples of 8 wide");
  if (tile == 0) {
    const long tiles128 = (long)((g.R + 127) / 128) * ((g.Cn + 127) / 128);
    tile = tiles128 >= 512 ? 4 : 1;
  }
  return prec == DENS_PREC_SPLIT_BF16 ? launch_gemm_bf_p<true>(g, g1, s, tile) : launch_gemm_bf_p<false>(g, g1, s, tile);
}

int launch_gemm(int prec, const Gemm& g, hipStream_t s, int tile) {
  if (g.K % 4 || g.lda % 4 || g.ldb % 4) return fail(DENS_ERR_ARG, "GEMM operands must be multiples of 4 wide");
  switch (prec) {
    case DENS_PREC_BF16: return launch_gemm_p<DENS_PREC_BF16>(g, s, tile);
    case DENS_PREC_SPLIT_BF16: return launch_gemm_p<DENS_PREC_SPLIT_BF16>(g, s, tile);
    case DENS_PREC_F32: return launch_gemm_p<DENS_PREC_F32>(g, s, tile);
  }
  return fail(DENS_ERR_ARG, "unknown precision");
}

// ---------------------------------------------------------------------------
// scatter / resolve / gather
// ---------------------------------------------------------------------------
struct Geo {
  int nbx, nby, tiles_x, T, H, SW, N;
  float die_lx, die_ly, bin_w, bin_h;
  const float *bin_lx, *bin_ux, *bin_ly, *bin_uy;
  // fixed-point scale of the bin sums: an overlap area a is added as the
  // integer rint(a * fx_scale) (order-independent, so every evaluation gives
  // the same bits), read back as sum * fx_inv
  float fx_scale;
  double fx_inv;
};

__device__ __forceinline__ unsigned long long to_fixed(const Geo& G, float a) {
  return (unsigned long long)__float2ll_rn(a * G.fx_scale);
}

// findBinWithDensitySize (DensityGradient.cu:20-59)
__device__ __forceinline__ void bin_range(const Geo& G, float cx, float cy, float dw, float dh, int& x0,
                                          int& x1, int& y0, int& y1) {
  float lx = cx - dw / 2, ux = cx + dw / 2;
  float ly = cy - dh / 2, uy = cy + dh / 2;
  x0 = max((int)floorf((lx - G.die_lx) / G.bin_w), 0);
  x1 = min((int)ceilf((ux - G.die_lx) / G.bin_w), G.nbx);
  y0 = max((int)floorf((ly - G.die_ly) / G.bin_h), 0);
  y1 = min((int)ceilf((uy - G.die_ly) / G.bin_h), G.nby);
}

// Overlap of the cell span [c - hw, c + hw] with the bin [b0, b1], measured
// from the bin's lower edge: c - b0 and b1 - b0 are exact (Sterbenz) for a
// cell near the bin, so the result keeps FP32 precision relative to the bin
// width instead of the absolute coordinate (on a 2500-unit die, the edges
// c -/+ hw alone would carry rounding errors of 1e-4 units per side).
__device__ __forceinline__ float overlap(float c, float hw, float b0, float b1) {
  const float d = c - b0;
  return fmaxf(0.f, fminf(d + hw, b1 - b0) - fmaxf(d - hw, 0.f));
}

__global__ void k_dens_reset(int* cnt, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) cnt[i] = 0;
}

// One rank per active lane from a counter per bucket, the lanes of a wave
// that share a bucket taking one atomic together (a clustered placement puts
// tens of thousands of cells in one tile, whose counter would otherwise take
// one atomic per cell).  Called by every lane of the wave.
template <typename Ctr>
__device__ __forceinline__ int wave_rank(Ctr* ctr, int b, bool active) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(active);
  int mine = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lb = __shfl(b, leader);
    const unsigned long long grp = __ballot(active && b == lb && ((pending >> lane) & 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[lb], (int)__popcll(grp));
    base = __shfl(base, leader);
    if ((grp >> lane) & 1ull) mine = base + (int)__popcll(grp & ((1ull << lane) - 1ull));
    pending &= ~grp;
  }
  return mine;
}

// bucket of each cell: its tile, T for cells wider than the halo, -1 for
// cells outside the die; rank within the bucket.  With HIST the ranks are
// taken per workgroup in an LDS histogram of the T + 1 buckets and the
// workgroup adds each bucket's total to the global count once.
template <bool HIST>
__global__ __launch_bounds__(256) void k_dens_count(Geo G, const float* cx, const float* cy, const float4* cgeo,
                                                    int* cnt, int* bucket, int* rank) {
  extern __shared__ int hist[];  // [T + 1] with HIST
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (HIST) {
    for (int e = threadIdx.x; e <= G.T; e += blockDim.x) hist[e] = 0;
    __syncthreads();
  }
  int b = -1;
  if (i < G.N) {
    float4 c = cgeo[i];
    int x0, x1, y0, y1;
    bin_range(G, cx[i], cy[i], c.x, c.y, x0, x1, y0, y1);
    if (x1 > x0 && y1 > y0) {
      bool fits = (x1 - x0) <= G.H + 1 && (y1 - y0) <= G.H + 1;
      b = fits ? (x0 / kTile) + (y0 / kTile) * G.tiles_x : G.T;
    }
    bucket[i] = b;
  }
  if (!HIST) {
    const int r = wave_rank(cnt, b, b >= 0);
    if (b >= 0) rank[i] = r;
    return;
  }
  const int rl = wave_rank(hist, b, b >= 0);
  __syncthreads();
  // the first cell of a bucket turns the bucket's total into its base
  if (b >= 0 && rl == 0) hist[b] = atomicAdd(&cnt[b], hist[b]);
  __syncthreads();
  if (b >= 0) rank[i] = hist[b] + rl;
}

// exclusive scan of one 1024-thread workgroup over n values given by f
template <typename F>
__device__ void block_scan(int n, F f, int* out, int* part) {
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(b + per, n);
  int s = 0;
  for (int i = b; i < e; i++) s += f(i);
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; i++) {
    out[i] = run;
    run += f(i);
  }
  if (t == 1023) out[n] = part[1023];
  __syncthreads();
}

// chunks of a tile's records (one workgroup each): at least one per tile, so
// that every tile's slab is written
__device__ __forceinline__ int tile_chunks(int c) { return c > kChunk ? (c + kChunk - 1) / kChunk : 1; }

// off[0..T+1] = exclusive scan of the bucket counts; cstart[0..T] = first
// chunk of each tile; mstart[0..T] = first partial slot of each tile that
// has more than one chunk
__global__ __launch_bounds__(1024) void k_dens_scan(const int* cnt, int* off, int T, int* cstart, int* mstart) {
  __shared__ int part[1024];
  block_scan(T + 1, [&](int i) { return cnt[i]; }, off, part);
  block_scan(T, [&](int i) { return tile_chunks(cnt[i]); }, cstart, part);
  block_scan(T, [&](int i) { int k = tile_chunks(cnt[i]); return k > 1 ? k : 0; }, mstart, part);
}

// chunk -> tile map
__global__ __launch_bounds__(256) void k_dens_chunkmap(int T, const int* cstart, int* cmap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  for (int c = cstart[t]; c < cstart[t + 1]; c++) cmap[c] = t;
}

// records in bucket order: box, scatter scale and filler flag, and for the
// gather the cell id and its gather scale; a cell outside the die gets its
// zero gradient here (it has no record)
__global__ void k_dens_fill(Geo G, const float* cx, const float* cy, const float4* cgeo, const int* cfill,
                            const int* bucket, const int* rank, const int* off, float4* rec_box, float2* rec_s,
                            int2* rec_ig, float* gx, float* gy, float* precond) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.N) return;
  int b = bucket[i];
  if (b < 0) {
    gx[i] = 0.f;
    gy[i] = 0.f;
    precond[i] = 0.f;
    return;
  }
  float4 c = cgeo[i];
  float x = cx[i], y = cy[i];
  int p = off[b] + rank[i];
  rec_box[p] = make_float4(x, y, c.x, c.y);
  rec_s[p] = make_float2(c.z, cfill ? (float)cfill[i] : 0.f);
  rec_ig[p] = make_int2(i, __float_as_int(c.w));
}

// one workgroup per chunk of at most kChunk records of one tile: LDS
// accumulation of the tile plus halo in fixed point (64-bit integer adds,
// whose order does not change the sum).  A tile of one chunk writes its slab
// in FP32; the chunks of a larger tile write their integer partials, which
// k_dens_merge adds (in any grouping of the tile's cells into chunks the
// integer total is the same).
__global__ __launch_bounds__(256) void k_dens_accum(Geo G, const int* off, const int* cstart, const int* mstart,
                                                    const int* cmap, const float4* rec_box, const float2* rec_s,
                                                    float* slab, unsigned long long* partial) {
  extern __shared__ unsigned long long sm[];  // [2][SW*SW]
  const int chunk = blockIdx.x;
  if (chunk >= cstart[G.T]) return;
  const int t = cmap[chunk];
  const int SW = G.SW, S2 = SW * SW;
  for (int e = threadIdx.x; e < 2 * S2; e += blockDim.x) sm[e] = 0ull;
  __syncthreads();
  const int tx0 = (t % G.tiles_x) * kTile, ty0 = (t / G.tiles_x) * kTile;
  const int k = chunk - cstart[t];
  const int b = off[t] + k * kChunk, e = min(off[t + 1], b + kChunk);
  for (int p = b + threadIdx.x; p < e; p += blockDim.x) {
    float4 c = rec_box[p];
    float2 s = rec_s[p];
    int x0, x1, y0, y1;
    bin_range(G, c.x, c.y, c.z, c.w, x0, x1, y0, y1);
    const float hw = c.z / 2, hh = c.w / 2;
    unsigned long long* dst = sm + (s.y != 0.f ? S2 : 0);
    for (int bx = x0; bx < x1; bx++) {
      float ox = overlap(c.x, hw, G.bin_lx[bx], G.bin_ux[bx]);
      for (int by = y0; by < y1; by++) {
        float oy = overlap(c.y, hh, G.bin_ly[by], G.bin_uy[by]);
        atomicAdd(&dst[(by - ty0) * SW + (bx - tx0)], to_fixed(G, ox * oy * s.x));
      }
    }
  }
  __syncthreads();
  if (cstart[t + 1] - cstart[t] == 1) {
    float* out = slab + (size_t)t * 2 * S2;
    for (int q = threadIdx.x; q < 2 * S2; q += blockDim.x) out[q] = (float)((double)(long long)sm[q] * G.fx_inv);
  } else {
    unsigned long long* out = partial + (size_t)(mstart[t] + k) * 2 * S2;
    for (int q = threadIdx.x; q < 2 * S2; q += blockDim.x) out[q] = sm[q];
  }
}

// the tiles of more than one chunk: integer sum of the chunks' partials, slab
// write in FP32
__global__ __launch_bounds__(256) void k_dens_merge(Geo G, const int* cstart, const int* mstart,
                                                    const unsigned long long* partial, float* slab) {
  const int t = blockIdx.x;
  const int nch = cstart[t + 1] - cstart[t];
  if (nch < 2) return;
  const int S2 = G.SW * G.SW;
  const unsigned long long* src = partial + (size_t)mstart[t] * 2 * S2;
  float* out = slab + (size_t)t * 2 * S2;
  for (int q = threadIdx.x; q < 2 * S2; q += blockDim.x) {
    unsigned long long v = 0ull;
    for (int c = 0; c < nch; c++) v += src[(size_t)c * 2 * S2 + q];
    out[q] = (float)((double)(long long)v * G.fx_inv);
  }
}

// cells wider than the halo: global 64-bit fixed-point atomics into big[2][B]
// (movable, filler), which the resolve pass adds and clears
__global__ __launch_bounds__(256) void k_dens_large(Geo G, const int* off, const float4* rec_box,
                                                    const float2* rec_s, unsigned long long* big) {
  const int b = off[G.T], e = off[G.T + 1];
  const size_t B = (size_t)G.nbx * G.nby;
  for (int p = b + blockIdx.x; p < e; p += gridDim.x) {
    float4 c = rec_box[p];
    float2 s = rec_s[p];
    int x0, x1, y0, y1;
    bin_range(G, c.x, c.y, c.z, c.w, x0, x1, y0, y1);
    const float hw = c.z / 2, hh = c.w / 2;
    const int nx = x1 - x0, cells = nx * (y1 - y0);
    for (int q = threadIdx.x; q < cells; q += blockDim.x) {
      int bx = x0 + q % nx, by = y0 + q / nx;
      float v = overlap(c.x, hw, G.bin_lx[bx], G.bin_ux[bx]) * overlap(c.y, hh, G.bin_ly[by], G.bin_uy[by]) * s.x;
      atomicAdd(&big[(s.y != 0.f ? B : 0) + (size_t)bx + (size_t)by * G.nbx], to_fixed(G, v));
    }
  }
}

// per bin: own slab entry plus the halo entries of the tiles left / below /
// below-left; density and overflow (computeOverflow, DensityGradient.cu:140-156)
__global__ __launch_bounds__(256) void k_dens_resolve(Geo G, const int* off, const float* slab,
                                                      unsigned long long* big, const float* fixed,
                                                      const float* scaled, float* mov, float* fil, float* dens,
                                                      double* ovf_part) {
  const int B = G.nbx * G.nby;
  const bool large = off[G.T + 1] > off[G.T];
  const int SW = G.SW, S2 = SW * SW;
  double ovf = 0.0;
  for (int bin = blockIdx.x * blockDim.x + threadIdx.x; bin < B; bin += gridDim.x * blockDim.x) {
    int x = bin % G.nbx, y = bin / G.nbx;
    int tx = x / kTile, ty = y / kTile, lx = x - tx * kTile, ly = y - ty * kTile;
    const float* s0 = slab + (size_t)(tx + ty * G.tiles_x) * 2 * S2;
    float m = s0[ly * SW + lx], f = s0[S2 + ly * SW + lx];
    if (lx < G.H && tx > 0) {
      const float* s1 = s0 - 2 * S2;
      m += s1[ly * SW + lx + kTile];
      f += s1[S2 + ly * SW + lx + kTile];
    }
    if (ly < G.H && ty > 0) {
      const float* s2 = s0 - (size_t)G.tiles_x * 2 * S2;
      m += s2[(ly + kTile) * SW + lx];
      f += s2[S2 + (ly + kTile) * SW + lx];
      if (lx < G.H && tx > 0) {
        const float* s3 = s2 - 2 * S2;
        m += s3[(ly + kTile) * SW + lx + kTile];
        f += s3[S2 + (ly + kTile) * SW + lx + kTile];
      }
    }
    if (large) {
      m += (float)((double)(long long)big[bin] * G.fx_inv);
      f += (float)((double)(long long)big[(size_t)B + bin] * G.fx_inv);
      big[bin] = 0ull;
      big[(size_t)B + bin] = 0ull;
    }
    mov[bin] = m;
    fil[bin] = f;
    float fx = fixed[bin], sc = scaled[bin];
    dens[bin] = (m + fx + f) / sc;
    ovf += (double)fmaxf(0.f, m + fx - sc);
  }
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) ovf += __shfl_down(ovf, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ovf;
  __syncthreads();
  if (threadIdx.x == 0) ovf_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// getDensityGradForEachCell (DensityGradient.cu:173-281): the GPU Poisson
// field carries a factor 2, hence the 0.5
// over the records in bucket (tile) order, so that neighbouring threads read
// neighbouring bins; results go to each record's cell
__global__ void k_dens_gather(Geo G, const int* off, const float4* rec_box, const int2* rec_ig, const float* lambda,
                              const float* ex, const float* ey, float* gx, float* gy, float* precond) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= off[G.T + 1]) return;
  const float4 r = rec_box[p];
  const int2 ig = rec_ig[p];
  const int i = ig.x;
  const float4 c = make_float4(r.z, r.w, 0.f, __int_as_float(ig.y));
  float x = r.x, y = r.y;
  int x0, x1, y0, y1;
  bin_range(G, x, y, c.x, c.y, x0, x1, y0, y1);
  const float hw = c.x / 2, hh = c.y / 2;
  float fx = 0.f, fy = 0.f, pc = 0.f;
  for (int bx = x0; bx < x1; bx++) {
    float ox = overlap(x, hw, G.bin_lx[bx], G.bin_ux[bx]);
    for (int by = y0; by < y1; by++) {
      float a = ox * overlap(y, hh, G.bin_ly[by], G.bin_uy[by]) * c.w;
      int b = bx + by * G.nbx;
      float la = lambda[b] * a;
      fx += la * ex[b] * 0.5f;
      fy += la * ey[b] * 0.5f;
      pc += la;
    }
  }
  gx[i] = fx;
  gy[i] = fy;
  precond[i] = pc;
}

// the overflow ratio from the resolve pass's per-block partials, summed in
// block order
__global__ __launch_bounds__(64) void k_dens_ratio(const double* ovf_part, int nparts, double inv_area, float* out) {
  __shared__ double sh[64];
  double s = 0.0;
  for (int k = threadIdx.x; k < nparts; k += 64) s += ovf_part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 64; k++) t += sh[k];
    *out = (float)(t * inv_area);
  }
}

template <typename T>
int dalloc(std::vector<void*>& bufs, T** p, size_t n) {
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return fail(DENS_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  bufs.push_back((void*)*p);
  return DENS_OK;
}

template <typename T>
int dupload(std::vector<void*>& bufs, T** p, const T* src, size_t n) {
  int st = dalloc(bufs, p, n);
  if (st) return st;
  if (n) DENS_TRY(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return DENS_OK;
}

}  // namespace densdev

using namespace densdev;

struct PoissonHandle {
  int M = 0, N = 0;
  int tile_mode = 0;                // 0: tuned per precision at first use, 1..4: fixed shape
  int tile[3] = {0, 0, 0};          // chosen shape per precision (0: not tuned yet)
  float tune_ms[3][kNumTiles] = {};  // measured chain time per shape
  std::vector<void*> bufs;
  float *CN = nullptr, *CM = nullptr, *A3 = nullptr, *A4 = nullptr, *wx2 = nullptr, *wy2 = nullptr;
  float *U = nullptr, *Ap = nullptr, *V = nullptr;
  // the tables split into bf16 hi / lo once (bin counts multiples of 8;
  // else the BF16 chains take the FP32-operand kernel)
  bool bf = false;
  __bf16 *CNh = nullptr, *CNl = nullptr, *CMh = nullptr, *CMl = nullptr;
  __bf16 *A3h = nullptr, *A3l = nullptr, *A4h = nullptr, *A4l = nullptr;
};

struct DensHandle {
  Geo G;
  int B = 0;
  int prec = DENS_PREC_SPLIT_BF16;
  double inv_sum_area = 0.0;
  std::vector<void*> bufs;
  float4* cgeo = nullptr;
  int* cfill = nullptr;
  float *fixed = nullptr, *scaled = nullptr, *lambda = nullptr;
  int *cnt = nullptr, *off = nullptr, *bucket = nullptr, *rank = nullptr;
  int *cstart = nullptr, *mstart = nullptr, *cmap = nullptr;  // chunks of the tiles' records
  unsigned long long* partial = nullptr;                      // [max multi-chunk chunks][2][SW*SW]
  int max_chunks = 0;
  float4* rec_box = nullptr;
  float2* rec_s = nullptr;
  int2* rec_ig = nullptr;
  float *slab = nullptr, *mov = nullptr, *fil = nullptr, *dens = nullptr;
  float *psi = nullptr, *ex = nullptr, *ey = nullptr, *precond = nullptr;
  unsigned long long* big = nullptr;  // [2][B] fixed-point sums of the cells wider than the halo
  double* ovf_part = nullptr;         // [kResolveBlocks]
  float* ovf_val = nullptr;
  PoissonHandle* ps = nullptr;
};

namespace {

// cosine / sine synthesis tables (FP64 on the host, stored FP32)
void poisson_tables(int M, int N, std::vector<float>& CN, std::vector<float>& CM, std::vector<float>& A3,
                    std::vector<float>& A4, std::vector<float>& wx2, std::vector<float>& wy2) {
  const double pi = 3.14159265358979323846;
  CN.assign((size_t)N * N, 0.f);
  CM.assign((size_t)M * M, 0.f);
  A3.assign((size_t)2 * N * N, 0.f);
  A4.assign((size_t)2 * M * M, 0.f);
  for (int j = 0; j < N; j++)
    for (int n = 0; n < N; n++) CN[(size_t)j * N + n] = (float)std::cos(pi * j * (2 * n + 1) / (2.0 * N));
  for (int u = 0; u < M; u++)
    for (int m = 0; m < M; m++) CM[(size_t)u * M + m] = (float)std::cos(pi * u * (2 * m + 1) / (2.0 * M));
  for (int n = 0; n < N; n++)
    for (int w = 0; w < N; w++) {
      double ang = pi * w * (2 * n + 1) / (2.0 * N);
      A3[(size_t)n * N + w] = (float)((w ? 2.0 : 1.0) * std::cos(ang));
      A3[(size_t)(N + n) * N + w] = w ? (float)(2.0 * std::sin(ang) * (2.0 * pi * w / N)) : 0.f;
    }
  for (int m = 0; m < M; m++)
    for (int h = 0; h < M; h++) {
      double ang = pi * h * (2 * m + 1) / (2.0 * M);
      A4[(size_t)m * M + h] = (float)((h ? 2.0 : 1.0) * std::cos(ang));
      A4[(size_t)(M + m) * M + h] = h ? (float)(2.0 * std::sin(ang) * (2.0 * pi * h / M)) : 0.f;
    }
  // divideByWSquare (PoissonSolver.cu:73-95): w = 2 pi idx / numBin, squared in FP32
  wx2.resize(N);
  wy2.resize(M);
  for (int w = 0; w < N; w++) {
    float d = (float)(2.0 * pi * w / N);
    wx2[w] = d * d;
  }
  for (int h = 0; h < M; h++) {
    float d = (float)(2.0 * pi * h / M);
    wy2[h] = d * d;
  }
}

// workspace of one solve: U [B], A' [B], [V; Vx] [2B]
struct PoissonWork {
  float *U, *Ap, *V;
};

int poisson_chain(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int prec, hipStream_t s,
                  int tile, PoissonWork w);

PoissonWork own_work(PoissonHandle* h) { return PoissonWork{h->U, h->Ap, h->V}; }

// Autotuner: the first solve of a precision times the whole 5-GEMM chain with
// every tile shape on the caller's own buffers (two runs each, the second
// timed with HIP events) and keeps the fastest.
int poisson_tune(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int prec, hipStream_t s) {
  hipEvent_t e0, e1;
  DENS_TRY(hipEventCreate(&e0));
  DENS_TRY(hipEventCreate(&e1));
  int best = 0, st = DENS_OK;
  float best_ms = 0.f;
  for (int t = 1; t <= kNumTiles && !st; t++) {
    if ((st = poisson_chain(h, rho, psi, ex, ey, prec, s, t, own_work(h)))) break;
    if (hipEventRecord(e0, s) != hipSuccess) st = fail(DENS_ERR_DEVICE, "hipEventRecord");
    if (!st) st = poisson_chain(h, rho, psi, ex, ey, prec, s, t, own_work(h));
    if (!st && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
      st = fail(DENS_ERR_DEVICE, "tuning events");
    float ms = 0.f;
    if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = fail(DENS_ERR_DEVICE, "hipEventElapsedTime");
    h->tune_ms[prec][t - 1] = ms;
    if (!st && (best == 0 || ms < best_ms)) {
      best = t;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!st) h->tile[prec] = best;
  return st;
}

int poisson_run(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int prec, hipStream_t s) {
  if (prec < 0 || prec > 2) return fail(DENS_ERR_ARG, "unknown precision");
  int tile = h->tile_mode;
  if (tile == 0) {
    if (h->tile[prec] == 0) {
      int st = poisson_tune(h, rho, psi, ex, ey, prec, s);
      if (st) return st;
    }
    tile = h->tile[prec];
  }
  return poisson_chain(h, rho, psi, ex, ey, prec, s, tile, own_work(h));
}

// ---- the allocation-free solve's workspace -------------------------------
// One static array per GPU the library runs on (held until the process
// ends), shared by every poisson_solve_noalloc call on that GPU.  The host
// name `scratch` is only a stand-in: kernels get the device address that
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocFloats = (size_t)4 << 20;  // 16 MiB: 4 * M * N <= 4M, e.g. 1024 x 1024 bins
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocFloats];
float* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(float** out) {
  int dev = 0;
  DENS_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DENS_ERR_DEVICE, "device index beyond the library's table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DENS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = (float*)p;
  }
  *out = g_scratch_dev[dev];
  return DENS_OK;
}

// The chain on bf16 operands.  The workspace (4 B floats) holds the same
// intermediates as bf16 hi / lo pairs: U hi, lo in the first B floats' bytes,
// A' in the next, [V; Vx] in the last 2 B; the input density's hi / lo sit in
// A''s place until GEMM 2 overwrites them.
int poisson_chain_bf(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int prec, hipStream_t s,
                     int tile, PoissonWork w) {
  const int M = h->M, N = h->N;
  const size_t B = (size_t)M * N;
  const bool split = prec == DENS_PREC_SPLIT_BF16;
  __bf16* Uh = reinterpret_cast<__bf16*>(w.U);
  __bf16* Ul = Uh + B;
  __bf16* Ph = reinterpret_cast<__bf16*>(w.Ap);
  __bf16* Pl = Ph + B;
  __bf16* Vh = reinterpret_cast<__bf16*>(w.V);
  __bf16* Vl = Vh + 2 * B;
  const int sb = (int)std::min<size_t>((B + 255) / 256, 4096);
  if (split) hipLaunchKernelGGL((k_split_bf16<true>), dim3(sb), dim3(256), 0, s, rho, Ph, Pl, B);
  else hipLaunchKernelGGL((k_split_bf16<false>), dim3(sb), dim3(256), 0, s, rho, Ph, Pl, B);
  DENS_TRY(hipGetLastError());
  int st;
  GemmBf g{};
  // U = CN * rho^T
  g = GemmBf{h->CNh, h->CNl, Ph, Pl, nullptr, nullptr, Uh, Ul, N, N, M, N, N, M, N, 1.f, nullptr, nullptr};
  if ((st = launch_gemm_bf(prec, g, s, tile))) return st;
  // A' = 4/(MN) CM * U^T / (wy^2 + wx^2)
  g = GemmBf{h->CMh, h->CMl, Uh, Ul, nullptr, nullptr, Ph, Pl, M, M, N, M, M, N, M,
             (float)(4.0 / ((double)M * N)), h->wy2, h->wx2};
  if ((st = launch_gemm_bf(prec, g, s, tile))) return st;
  // [V; Vx] = [DNt; SNw] * A'^T
  g = GemmBf{h->A3h, h->A3l, Ph, Pl, nullptr, nullptr, Vh, Vl, N, N, M, 2 * N, 2 * N, M, N, 1.f, nullptr, nullptr};
  if ((st = launch_gemm_bf(prec, g, s, tile))) return st;
  // [psi; Ey] = [DMt; SMw] * V^T and Ex = DMt * Vx^T: both read A4 and
  // [V; Vx] only, so they share one launch
  g = GemmBf{h->A4h, h->A4l, Vh, Vl, psi, ey, nullptr, nullptr, M, M, N, M, 2 * M, N, M, 1.f, nullptr, nullptr};
  const GemmBf gx{h->A4h, h->A4l, Vh + B, Vl + B, ex, nullptr, nullptr, nullptr, M, M, N, M, M, N, M, 1.f, nullptr,
                  nullptr};
  return launch_gemm_bf(prec, g, s, tile, &gx);
}

// Every element of U, A' and [V; Vx] is written before it is read, so the
// workspace needs no clearing between solves.
int poisson_chain(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int prec, hipStream_t s,
                  int tile, PoissonWork w) {
  const int M = h->M, N = h->N;
  if (h->bf && prec != DENS_PREC_F32) return poisson_chain_bf(h, rho, psi, ex, ey, prec, s, tile, w);
  int st;
  Gemm g{};
  // U = CN * rho^T
  g = Gemm{h->CN, rho, w.U, nullptr, N, N, M, N, N, M, N, 1.f, nullptr, nullptr};
  if ((st = launch_gemm(prec, g, s, tile))) return st;
  // A' = 4/(MN) CM * U^T / (wy^2 + wx^2)
  g = Gemm{h->CM, w.U, w.Ap, nullptr, M, M, N, M, M, N, M, (float)(4.0 / ((double)M * N)), h->wy2, h->wx2};
  if ((st = launch_gemm(prec, g, s, tile))) return st;
  // [V; Vx] = [DNt; SNw] * A'^T
  g = Gemm{h->A3, w.Ap, w.V, nullptr, N, N, M, 2 * N, 2 * N, M, N, 1.f, nullptr, nullptr};
  if ((st = launch_gemm(prec, g, s, tile))) return st;
  // [psi; Ey] = [DMt; SMw] * V^T
  g = Gemm{h->A4, w.V, psi, ey, M, M, N, M, 2 * M, N, M, 1.f, nullptr, nullptr};
  if ((st = launch_gemm(prec, g, s, tile))) return st;
  // Ex = DMt * Vx^T
  g = Gemm{h->A4, w.V + (size_t)N * M, ex, nullptr, M, M, N, M, M, N, M, 1.f, nullptr, nullptr};
  return launch_gemm(prec, g, s, tile);
}


}  // namespace

extern "C" {

const char* dens_last_error(void) { return g_err.c_str(); }

int poisson_create(int32_t nbx, int32_t nby, PoissonHandle** out) {
  if (!out) return fail(DENS_ERR_ARG, "null argument");
  if (nbx < 4 || nby < 4 || nbx % 4 || nby % 4 || nbx > 16384 || nby > 16384)
    return fail(DENS_ERR_ARG, "bin counts must be multiples of 4 in [4, 16384]");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DENS_ERR_DEVICE, "no HIP device");
  PoissonHandle* h = new PoissonHandle();
  h->M = nby;
  h->N = nbx;
  std::vector<float> CN, CM, A3, A4, wx2, wy2;
  poisson_tables(nby, nbx, CN, CM, A3, A4, wx2, wy2);
  const size_t B = (size_t)nbx * nby;
  int st;
  auto bail = [&](int c) { poisson_destroy(h); return c; };
  if ((st = dupload(h->bufs, &h->CN, CN.data(), CN.size()))) return bail(st);
  if ((st = dupload(h->bufs, &h->CM, CM.data(), CM.size()))) return bail(st);
  if ((st = dupload(h->bufs, &h->A3, A3.data(), A3.size()))) return bail(st);
  if ((st = dupload(h->bufs, &h->A4, A4.data(), A4.size()))) return bail(st);
  if ((st = dupload(h->bufs, &h->wx2, wx2.data(), wx2.size()))) return bail(st);
  if ((st = dupload(h->bufs, &h->wy2, wy2.data(), wy2.size()))) return bail(st);
  if ((st = dalloc(h->bufs, &h->U, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->Ap, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->V, 2 * B))) return bail(st);
  // SKYLINE_POISSON_FP32_OPERANDS=1 keeps the FP32-operand kernel for every
  // precision (the two give the same bits; tests compare them)
  const char* fp32_ops = std::getenv("SKYLINE_POISSON_FP32_OPERANDS");
  if (nbx % 8 == 0 && nby % 8 == 0 && !(fp32_ops && fp32_ops[0] == '1')) {
    // the tables' bf16 hi / lo, split on the device as the kernels split
    struct Part { const float* src; size_t n; __bf16** hi; __bf16** lo; };
    const Part parts[4] = {{h->CN, CN.size(), &h->CNh, &h->CNl}, {h->CM, CM.size(), &h->CMh, &h->CMl},
                           {h->A3, A3.size(), &h->A3h, &h->A3l}, {h->A4, A4.size(), &h->A4h, &h->A4l}};
    for (const Part& q : parts) {
      __bf16* d = nullptr;
      if ((st = dalloc(h->bufs, &d, 2 * q.n))) return bail(st);
      *q.hi = d;
      *q.lo = d + q.n;
      hipLaunchKernelGGL((k_split_bf16<true>), dim3(1024), dim3(256), 0, 0, q.src, *q.hi, *q.lo, q.n);
    }
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(DENS_ERR_DEVICE, "splitting the tables"));
    h->bf = true;
  }
  *out = h;
  return DENS_OK;
}

int poisson_solve(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int32_t prec,
                  void* stream) {
  if (!h || !rho || !psi || !ex || !ey) return fail(DENS_ERR_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  return poisson_run(h, rho, psi, ex, ey, prec, s);
}

int poisson_solve_noalloc(PoissonHandle* h, const float* rho, float* psi, float* ex, float* ey, int32_t prec,
                          void* stream) {
  if (!h || !rho || !psi || !ex || !ey) return fail(DENS_ERR_ARG, "null argument");
  if (prec < 0 || prec > 2) return fail(DENS_ERR_ARG, "unknown precision");
  const size_t B = (size_t)h->M * h->N;
  if (4 * B > kNoAllocFloats)
    return fail(DENS_ERR_ARG, "poisson_solve_noalloc: " + std::to_string(h->N) + " x " + std::to_string(h->M) +
                                  " bins need " + std::to_string(16 * B) + " bytes of workspace, more than its " +
                                  std::to_string(4 * kNoAllocFloats) + "-byte static array");
  float* w = nullptr;
  int st = scratch_for_current_device(&w);
  if (st) return st;
  // the tile of poisson_solve without tuning: fixed, or tuned before, or
  // 64 x 64; every shape gives the same bits
  int tile = h->tile_mode ? h->tile_mode : (h->tile[prec] ? h->tile[prec] : 1);
  return poisson_chain(h, rho, psi, ex, ey, prec, (hipStream_t)stream, tile, PoissonWork{w, w + B, w + 2 * B});
}

int poisson_set_tile(PoissonHandle* h, int32_t tile) {
  if (!h || tile < 0 || tile > kNumTiles) return fail(DENS_ERR_ARG, "tile must be 0 (tuned) or 1..5");
  h->tile_mode = tile;
  return DENS_OK;
}

int poisson_get_tile(PoissonHandle* h, int32_t prec, int32_t* tile, float* ms) {
  if (!h || prec < 0 || prec > 2) return fail(DENS_ERR_ARG, "bad argument");
  if (tile) *tile = h->tile_mode ? h->tile_mode : h->tile[prec];
  if (ms)
    for (int t = 0; t < kNumTiles; t++) ms[t] = h->tune_ms[prec][t];
  return DENS_OK;
}

int poisson_destroy(PoissonHandle* h) {
  if (!h) return DENS_OK;
  for (void* b : h->bufs) (void)hipFree(b);
  delete h;
  return DENS_OK;
}

int dens_create(const DensGeom* g, DensHandle** out) {
  if (!g || !out) return fail(DENS_ERR_ARG, "null argument");
  if (g->num_movable < 0 || !g->bin_lx || !g->bin_ux || !g->bin_ly || !g->bin_uy || !g->fixed_area ||
      !g->scaled_bin_area)
    return fail(DENS_ERR_ARG, "missing geometry arrays");
  if (g->num_movable > 0 && (!g->cell_dw || !g->cell_dh || !g->cell_scale))
    return fail(DENS_ERR_ARG, "missing cell arrays");
  if (!(g->bin_w > 0.f) || !(g->bin_h > 0.f)) return fail(DENS_ERR_ARG, "bad bin size");
  if (g->precision < 0 || g->precision > 2) return fail(DENS_ERR_ARG, "unknown precision");
  PoissonHandle* ps = nullptr;
  int st = poisson_create(g->num_bin_x, g->num_bin_y, &ps);
  if (st) return st;
  DensHandle* h = new DensHandle();
  h->ps = ps;
  auto bail = [&](int c) { dens_destroy(h); return c; };
  const int nbx = g->num_bin_x, nby = g->num_bin_y, N = g->num_movable;
  const size_t B = (size_t)nbx * nby;
  h->B = (int)B;
  h->prec = g->precision;
  h->inv_sum_area = g->sum_movable_area > 0.f ? 1.0 / (double)g->sum_movable_area : 0.0;
  // halo: widest density footprint in bins (a footprint of width w spans at
  // most ceil(w / bin) + 1 bins), capped; wider cells take the atomic path
  float maxw = 0.f, maxh = 0.f;
  std::vector<float4> cg(N);
  std::vector<int> cf(N, 0);
  for (int i = 0; i < N; i++) {
    float dw = g->cell_dw[i], dh = g->cell_dh[i], sc = g->cell_scale[i];
    bool mac = g->is_macro && g->is_macro[i];
    cg[i] = make_float4(dw, dh, mac ? g->target_density * sc : sc, sc);
    cf[i] = g->is_filler && g->is_filler[i] ? 1 : 0;
    maxw = std::max(maxw, dw);
    maxh = std::max(maxh, dh);
  }
  int H = (int)std::ceil(std::max(maxw / g->bin_w, maxh / g->bin_h));
  H = std::min(std::max(H, 1), kMaxHalo);
  Geo& G = h->G;
  G.nbx = nbx; G.nby = nby; G.N = N; G.H = H; G.SW = kTile + H;
  G.tiles_x = (nbx + kTile - 1) / kTile;
  G.T = G.tiles_x * ((nby + kTile - 1) / kTile);
  G.die_lx = g->die_lx; G.die_ly = g->die_ly; G.bin_w = g->bin_w; G.bin_h = g->bin_h;
  float *blx, *bux, *bly, *buy;
  if ((st = dupload(h->bufs, &blx, g->bin_lx, nbx))) return bail(st);
  if ((st = dupload(h->bufs, &bux, g->bin_ux, nbx))) return bail(st);
  if ((st = dupload(h->bufs, &bly, g->bin_ly, nby))) return bail(st);
  if ((st = dupload(h->bufs, &buy, g->bin_uy, nby))) return bail(st);
  G.bin_lx = blx; G.bin_ux = bux; G.bin_ly = bly; G.bin_uy = buy;
  // fixed-point scale: sums of up to 2^20 bin areas fit in 62 bits
  float max_bw = 0.f, max_bh = 0.f;
  for (int x = 0; x < nbx; x++) max_bw = std::max(max_bw, g->bin_ux[x] - g->bin_lx[x]);
  for (int y = 0; y < nby; y++) max_bh = std::max(max_bh, g->bin_uy[y] - g->bin_ly[y]);
  const double amax = std::max((double)max_bw * (double)max_bh * std::max(1.0f, g->target_density), 1e-30);
  const int fx_exp = std::min(100, (int)std::floor(62.0 - 20.0 - std::log2(amax)));
  G.fx_scale = std::ldexp(1.0f, fx_exp);
  G.fx_inv = std::ldexp(1.0, -fx_exp);
  if ((st = dupload(h->bufs, &h->cgeo, cg.data(), N))) return bail(st);
  if ((st = dupload(h->bufs, &h->cfill, cf.data(), N))) return bail(st);
  if ((st = dupload(h->bufs, &h->fixed, g->fixed_area, B))) return bail(st);
  if ((st = dupload(h->bufs, &h->scaled, g->scaled_bin_area, B))) return bail(st);
  std::vector<float> ones;
  const float* lam = g->bin_lambda;
  if (!lam) { ones.assign(B, 1.f); lam = ones.data(); }
  if ((st = dupload(h->bufs, &h->lambda, lam, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->cnt, G.T + 1))) return bail(st);
  if ((st = dalloc(h->bufs, &h->off, G.T + 2))) return bail(st);
  // chunks: at most one per tile plus N / kChunk; the partials of tiles of
  // more than one chunk: at most 2 N / kChunk chunks
  h->max_chunks = G.T + (N + kChunk - 1) / kChunk;
  const size_t max_multi = 2 * (size_t)((N + kChunk - 1) / kChunk);
  if ((st = dalloc(h->bufs, &h->cstart, G.T + 1))) return bail(st);
  if ((st = dalloc(h->bufs, &h->mstart, G.T + 1))) return bail(st);
  if ((st = dalloc(h->bufs, &h->cmap, h->max_chunks))) return bail(st);
  if ((st = dalloc(h->bufs, &h->partial, std::max<size_t>(max_multi, 1) * 2 * (size_t)G.SW * G.SW))) return bail(st);
  if ((st = dalloc(h->bufs, &h->bucket, N))) return bail(st);
  if ((st = dalloc(h->bufs, &h->rank, N))) return bail(st);
  if ((st = dalloc(h->bufs, &h->rec_box, N))) return bail(st);
  if ((st = dalloc(h->bufs, &h->rec_s, N))) return bail(st);
  if ((st = dalloc(h->bufs, &h->rec_ig, N))) return bail(st);
  if ((st = dalloc(h->bufs, &h->slab, (size_t)G.T * 2 * G.SW * G.SW))) return bail(st);
  if ((st = dalloc(h->bufs, &h->mov, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->fil, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->dens, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->psi, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->ex, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->ey, B))) return bail(st);
  if ((st = dalloc(h->bufs, &h->precond, N))) return bail(st);
  if ((st = dalloc(h->bufs, &h->big, 2 * B))) return bail(st);
  if (hipMemset(h->big, 0, 2 * B * sizeof(unsigned long long)) != hipSuccess)
    return bail(fail(DENS_ERR_DEVICE, "hipMemset"));
  if ((st = dalloc(h->bufs, &h->ovf_part, kResolveBlocks))) return bail(st);
  if ((st = dalloc(h->bufs, &h->ovf_val, 1))) return bail(st);
  // tile shape of the Poisson GEMMs tuned now for the handle's precision, so
  // that dens_compute_grad does not wait on the host (the timing does not
  // depend on the values); another precision is tuned at its first use
  if (hipMemset(h->dens, 0, (size_t)B * sizeof(float)) != hipSuccess)
    return bail(fail(DENS_ERR_DEVICE, "hipMemset"));
  if ((st = poisson_tune(h->ps, h->dens, h->psi, h->ex, h->ey, h->prec, nullptr))) return bail(st);
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(DENS_ERR_DEVICE, "tuning"));
  *out = h;
  return DENS_OK;
}

int dens_compute_grad(DensHandle* h, const float* cx, const float* cy, float* gx, float* gy, float* overflow_out,
                      void* stream) {
  if (!h) return fail(DENS_ERR_ARG, "null handle");
  const Geo& G = h->G;
  if (G.N > 0 && (!cx || !cy || !gx || !gy)) return fail(DENS_ERR_ARG, "null cell array");
  hipStream_t s = (hipStream_t)stream;
  const int nb = (G.N + 255) / 256;
  hipLaunchKernelGGL(k_dens_reset, dim3((G.T + 256) / 256), dim3(256), 0, s, h->cnt, G.T + 1);
  if (G.N > 0) {
    // LDS histogram of the buckets when it is small (up to 64 x 64 tiles)
    const size_t hist = (size_t)(G.T + 1) * sizeof(int);
    if (hist <= 16 * 1024 + sizeof(int))
      hipLaunchKernelGGL(k_dens_count<true>, dim3(nb), dim3(256), hist, s, G, cx, cy, h->cgeo, h->cnt, h->bucket,
                         h->rank);
    else
      hipLaunchKernelGGL(k_dens_count<false>, dim3(nb), dim3(256), 0, s, G, cx, cy, h->cgeo, h->cnt, h->bucket,
                         h->rank);
  }
  hipLaunchKernelGGL(k_dens_scan, dim3(1), dim3(1024), 0, s, h->cnt, h->off, G.T, h->cstart, h->mstart);
  hipLaunchKernelGGL(k_dens_chunkmap, dim3((G.T + 255) / 256), dim3(256), 0, s, G.T, h->cstart, h->cmap);
  if (G.N > 0) {
    hipLaunchKernelGGL(k_dens_fill, dim3(nb), dim3(256), 0, s, G, cx, cy, h->cgeo, h->cfill, h->bucket, h->rank,
                       h->off, h->rec_box, h->rec_s, h->rec_ig, gx, gy, h->precond);
  }
  size_t lds = (size_t)2 * G.SW * G.SW * sizeof(unsigned long long);
  hipLaunchKernelGGL(k_dens_accum, dim3(h->max_chunks), dim3(256), lds, s, G, h->off, h->cstart, h->mstart, h->cmap,
                     h->rec_box, h->rec_s, h->slab, h->partial);
  hipLaunchKernelGGL(k_dens_merge, dim3(G.T), dim3(256), 0, s, G, h->cstart, h->mstart, h->partial, h->slab);
  hipLaunchKernelGGL(k_dens_large, dim3(256), dim3(256), 0, s, G, h->off, h->rec_box, h->rec_s, h->big);
  const int nres = std::min((h->B + 255) / 256, kResolveBlocks);
  hipLaunchKernelGGL(k_dens_resolve, dim3(nres), dim3(256), 0, s, G, h->off, h->slab, h->big, h->fixed, h->scaled,
                     h->mov, h->fil, h->dens, h->ovf_part);
  hipLaunchKernelGGL(k_dens_ratio, dim3(1), dim3(64), 0, s, h->ovf_part, nres, h->inv_sum_area, h->ovf_val);
  DENS_TRY(hipGetLastError());
  int st = poisson_run(h->ps, h->dens, h->psi, h->ex, h->ey, h->prec, s);
  if (st) return st;
  if (G.N > 0) {
    hipLaunchKernelGGL(k_dens_gather, dim3(nb), dim3(256), 0, s, G, h->off, h->rec_box, h->rec_ig, h->lambda, h->ex,
                       h->ey, gx,
                       gy, h->precond);
  }
  DENS_TRY(hipGetLastError());
  if (overflow_out) {
    DENS_TRY(hipMemcpyAsync(overflow_out, h->ovf_val, sizeof(float), hipMemcpyDeviceToHost, s));
    DENS_TRY(hipStreamSynchronize(s));
  }
  return DENS_OK;
}

int dens_get_overflow(DensHandle* h, float* out) {
  if (!h || !out) return fail(DENS_ERR_ARG, "null argument");
  DENS_TRY(hipDeviceSynchronize());
  DENS_TRY(hipMemcpy(out, h->ovf_val, sizeof(float), hipMemcpyDeviceToHost));
  return DENS_OK;
}

const float* dens_get_overflow_device(DensHandle* h) { return h ? h->ovf_val : nullptr; }
const float* dens_get_bin_density(DensHandle* h) { return h ? h->dens : nullptr; }
const float* dens_get_potential(DensHandle* h) { return h ? h->psi : nullptr; }
const float* dens_get_field_x(DensHandle* h) { return h ? h->ex : nullptr; }
const float* dens_get_field_y(DensHandle* h) { return h ? h->ey : nullptr; }
const float* dens_get_preconditioner(DensHandle* h) { return h ? h->precond : nullptr; }
const float* dens_get_movable_area(DensHandle* h) { return h ? h->mov : nullptr; }
const float* dens_get_filler_area(DensHandle* h) { return h ? h->fil : nullptr; }

int dens_set_precision(DensHandle* h, int32_t prec) {
  if (!h) return fail(DENS_ERR_ARG, "null handle");
  if (prec < 0 || prec > 2) return fail(DENS_ERR_ARG, "unknown precision");
  h->prec = prec;
  return DENS_OK;
}

int dens_destroy(DensHandle* h) {
  if (!h) return DENS_OK;
  for (void* b : h->bufs) (void)hipFree(b);
  if (h->ps) poisson_destroy(h->ps);
  delete h;
  return DENS_OK;
}

}  // extern "C"
