/* C ABI of the MI355X electrostatic density path (SkyPlace DensityGradient /
 * PoissonSolver).
 *
 * Replaces, in skyplace/src/density_gradient/:
 *   - DensityGradient (DensityGradient.h:23-152): ctor / initForCUDAKernel
 *     (DensityGradient.cu:405-580) -> dens_create; computeGrad
 *     (DensityGradient.cu:302-403) -> dens_compute_grad; getters overflow(),
 *     getDevicePotential(), getDeviceBinDensity(), getDevicePreconditioner()
 *     -> dens_get_*;
 *   - PoissonSolver (PoissonSolver.h:9-68): ctor (PoissonSolver.cu:120-131) ->
 *     poisson_create; solvePoisson (PoissonSolver.cu:224-302) ->
 *     poisson_solve;
 *   - SkyPlaceDB::createBins / updateDensitySize / updateFixedOverlapArea
 *     (SkyPlaceDB.cpp:446-549, 821-852, 909-938) -> dens_prepare (host).
 *
 * Bins are y-major: bin = x + y * num_bin_x.  Cell coordinates are centres.
 * Device pointers are plain HIP device addresses; every call that takes a
 * stream enqueues on that hipStream_t (NULL is the HIP null stream) and
 * returns without synchronising unless stated.  Status: 0 = OK, else DENS_ERR_*; the text of
 * the last failure on this thread is in dens_last_error().
 */
#ifndef SKYLINE_DENSITY_H
#define SKYLINE_DENSITY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DENS_OK = 0,
  DENS_ERR_ARG = 1,
  DENS_ERR_DEVICE = 2,
  DENS_ERR_ALLOC = 3,
};

/* Precision of the Poisson transforms (cosine / sine GEMMs on MFMA, FP32
 * accumulation).  SPLIT_BF16 represents each FP32 operand as hi + lo bf16 and
 * sums three products (hi*hi + hi*lo + lo*hi). */
enum {
  DENS_PREC_BF16 = 0,
  DENS_PREC_SPLIT_BF16 = 1,
  DENS_PREC_F32 = 2,
};

/* Host description of a placement for dens_prepare. */
typedef struct DensDesign {
  float die_lx, die_ly, die_ux, die_uy;
  float target_density;
  int32_t num_movable;
  const float* cell_dx; /* movable cell sizes */
  const float* cell_dy;
  const int32_t* is_macro;  /* may be NULL */
  const int32_t* is_filler; /* may be NULL */
  int32_t num_fixed;
  const float* fixed_lx; /* fixed cell rectangles (lower-left, upper-right) */
  const float* fixed_ly;
  const float* fixed_ux;
  const float* fixed_uy;
  /* explicit bin counts (both > 0) override the automatic rule, which caps
   * at 1024 per side; 0 = automatic. */
  int32_t num_bin_x, num_bin_y;
} DensDesign;

/* Bin and cell geometry consumed by dens_create (host arrays). */
typedef struct DensGeom {
  int32_t num_bin_x, num_bin_y;
  float die_lx, die_ly, die_ux, die_uy;
  float bin_w, bin_h; /* nominal bin size (binX_, binY_) */
  float target_density;
  const float* bin_lx; /* [num_bin_x] */
  const float* bin_ux;
  const float* bin_ly; /* [num_bin_y] */
  const float* bin_uy;
  const float* fixed_area;      /* [B] fixed overlap area x target density */
  const float* scaled_bin_area; /* [B] bin area x target density */
  const float* bin_lambda;      /* [B] local lambda, NULL = all 1 */
  int32_t num_movable;
  const float* cell_dw; /* density width / height / scale [num_movable] */
  const float* cell_dh;
  const float* cell_scale;
  const int32_t* is_filler; /* may be NULL */
  const int32_t* is_macro;  /* may be NULL */
  float sum_movable_area;
  int32_t precision; /* DENS_PREC_* used by dens_compute_grad */
} DensGeom;

typedef struct DensHandle DensHandle;
typedef struct PoissonHandle PoissonHandle;

const char* dens_last_error(void);

/* SkyPlaceDB bin / density-size / fixed-area setup.  Cells flagged in
 * is_filler are sized like the others but left out of sum_movable_area and
 * of the automatic bin count, which the reference derives before
 * createFillers (SkyPlaceDB.cpp:320-355).  Call with
 * bin_lx == NULL to receive shape[0..1] = num_bin_x, num_bin_y; then with
 * arrays of that size: bin bounds, fixed_area and scaled_bin_area [B],
 * density sizes and scales [num_movable]; scalars[0..2] = bin_w, bin_h,
 * sum_movable_area. */
int dens_prepare(const DensDesign* d, int32_t* shape, float* bin_lx, float* bin_ux, float* bin_ly,
                 float* bin_uy, float* fixed_area, float* scaled_bin_area, float* cell_dw,
                 float* cell_dh, float* cell_scale, float* scalars);

/* Tunes the Poisson solver's tile shape for g->precision up front (so that
 * dens_compute_grad never waits on the host); another precision chosen with
 * dens_set_precision is tuned at its first evaluation, which then waits for
 * the stream once. */
int dens_create(const DensGeom* g, DensHandle** out);
/* One density evaluation (DensityGradient::computeGrad): scatter, overflow,
 * Poisson solve, gradient.  d_cx / d_cy are cell centres, d_gx / d_gy
 * receive the density gradient (N floats each).  If overflow_out is not NULL
 * the call synchronises the stream and stores the overflow ratio there. */
int dens_compute_grad(DensHandle* h, const float* d_cx, const float* d_cy, float* d_gx, float* d_gy,
                      float* overflow_out, void* stream);
/* Overflow of the last dens_compute_grad (synchronises). */
int dens_get_overflow(DensHandle* h, float* out);
/* Device address of the overflow ratio of the last dens_compute_grad (one
 * float, valid once the stream reaches it): lets a caller read it with other
 * values in one copy.  No counterpart in the reference (a host float there). */
const float* dens_get_overflow_device(DensHandle* h);
/* Device arrays owned by the handle: bin density, potential, field x / y
 * ([B] each) and the per-cell preconditioner ([N]). */
const float* dens_get_bin_density(DensHandle* h);
const float* dens_get_potential(DensHandle* h);
const float* dens_get_field_x(DensHandle* h);
const float* dens_get_field_y(DensHandle* h);
const float* dens_get_preconditioner(DensHandle* h);
/* Movable / filler area per bin of the last evaluation ([B] each). */
const float* dens_get_movable_area(DensHandle* h);
const float* dens_get_filler_area(DensHandle* h);
int dens_set_precision(DensHandle* h, int32_t precision);
int dens_destroy(DensHandle* h);

int poisson_create(int32_t num_bin_x, int32_t num_bin_y, PoissonHandle** out);
/* potential, field x, field y from bin density (all [B] device arrays,
 * y-major).  Unless a tile shape is fixed with poisson_set_tile, the first
 * solve of each precision is the autotuner's: it times the five GEMMs with
 * every tile shape on the given buffers, keeps the fastest (results do not
 * depend on the shape) and waits for the stream once.  dens_create tunes its
 * own solver for all three precisions up front. */
int poisson_solve(PoissonHandle* h, const float* d_rho, float* d_psi, float* d_ex, float* d_ey,
                  int32_t precision, void* stream);
/* poisson_solve without any device-memory allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * solve's workspace (4 * num_bin_x * num_bin_y floats) is a static array of
 * 16 MiB per GPU, held until the process ends: grids of more than 1024 x 1024
 * bins (4M floats) are rejected with DENS_ERR_ARG.  The tuner never runs
 * here: the tile shape is the fixed or already tuned one, else 64 x 64, so
 * the results are poisson_solve's bit for bit.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU race (serialise them, e.g.
 * on one stream).  Nothing in the array needs clearing between calls. */
int poisson_solve_noalloc(PoissonHandle* h, const float* d_rho, float* d_psi, float* d_ex, float* d_ey,
                          int32_t precision, void* stream);
/* GEMM tile shape: 0 = tuned per precision (default), 1 = 64x64, 2 = 64x128,
 * 3 = 128x64, 4 = 128x128, 5 = 256x128.  No counterpart in the reference
 * (cuFFT there). */
int poisson_set_tile(PoissonHandle* h, int32_t tile);
/* The shape in use for `precision` (0 before tuning) and, if ms != NULL, the
 * tuner's time of the whole solve per shape [5] (zeros if not tuned). */
int poisson_get_tile(PoissonHandle* h, int32_t precision, int32_t* tile, float* ms);
int poisson_destroy(PoissonHandle* h);

#ifdef __cplusplus
}
#endif

#endif /* SKYLINE_DENSITY_H */
