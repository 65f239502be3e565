/*
 * skyline_gp.h -- C ABI of the MI355X (gfx950) global placer that produces the
 * legalizer's input: SkyPlace's random initial placement and filler cells,
 * and the Nesterov (default) or Adam loop over the weighted-average
 * wirelength plus electrostatic density objective.
 *
 * Replaces, in src/skyline/skyplace/src/:
 *   - SkyPlace::run (SkyPlace.cpp:183-213) -> gp_place_design, or the handle
 *     sequence gp_create .. gp_solve .. gp_export;
 *   - SkyPlaceDB::init (SkyPlaceDB.cpp:320-355): importDB, createBins,
 *     createFillers (551-662), updateDensitySize, updateFixedOverlapArea ->
 *     gp_create;
 *   - InitialPlacer::doRandomInit (initial_place/InitialPlacer.cpp:77-123) ->
 *     gp_initial_place;
 *   - TargetFunction (problem_instance/TargetFunction.cu:17-299) and
 *     NesterovSolver (nonlinear_solver/NesterovSolver.cu:54-264) /
 *     AdamSolver (nonlinear_solver/AdamSolver.cu:38-226) -> gp_init_solver,
 *     gp_iterate, gp_solve;
 *   - SkyPlaceDB::exportDB (SkyPlaceDB.cpp:120-138) -> gp_export.
 *
 * The design is the legalizer's LgDesign (include/skyline_lg.h) with its
 * netlist; coordinates are dbu divided by GpParams.scaling_factor, as floats.  Movable cells are the instances that are not fixed, in
 * instance order, followed by the fillers.  A pin of a movable cell sits at
 * the cell centre plus (pin_dx, pin_dy), as GPPin::updatePinLocation places
 * it with the master pin's offset (object/GPPin.cpp:26-36, 61-68); a pin of
 * a fixed instance at that instance's centre plus the offset; an IO pin at
 * (pin_dx, pin_dy).
 *
 * Every entry point returns 0 (GP_OK) or an error code, with a message for
 * the calling thread in gp_last_error(); the library never exits.
 */
#ifndef SKYLINE_GP_H
#define SKYLINE_GP_H

#include <stdint.h>

#include "skyline_lg.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GP_OK = 0,
  GP_ERR_ARG = 1,     /* bad argument or design (e.g. target density <= utilisation) */
  GP_ERR_DEVICE = 2,  /* HIP runtime error */
  GP_ERR_MEMORY = 4,  /* device allocation failed */
  GP_ERR_STATE = 5    /* call out of order */
};

enum { GP_SOLVER_NESTEROV = 0, GP_SOLVER_ADAM = 1 };

/* HyperParam (HyperParam.h:23-48) and the SkyPlace / SkyPlaceDB setters.
 * gp_default_params fills the reference's defaults. */
typedef struct GpParams {
  int32_t max_opt_iter;        /* maxOptIter        2500  */
  int32_t max_backtrack_iter;  /* maxBackTrackIter  10    */
  float init_lambda;           /* initLambda        1e-4  */
  float init_gamma_inv_coef;   /* initGammaInvCoef  0.01  */
  float target_overflow;       /* targetOverflow    0.07  */
  float min_phi_coef;          /* minPhiCoef        0.95  */
  float max_phi_coef;          /* maxPhiCoef        1.05  */
  float min_precond;           /* minPrecond        1     */
  float init_opt_coef;         /* initOptCoef       100   */
  float reference_hpwl;        /* referenceHpwl     2e5   */
  float reference_hpwl_ratio;  /* 0: use reference_hpwl; > 0: reference_hpwl = ratio x the HPWL of
                                  the design's own positions (no counterpart in the reference, whose
                                  2e5 suits designs of its benchmarks' coordinate scale) */
  float min_step_length;       /* minStepLength     1     */
  float adam_alpha;            /* 100   */
  float adam_beta1;            /* 0.9   */
  float adam_beta2;            /* 0.999 */
  float adam_epsilon;          /* 1e-8 (the reference leaves AdamSolver::epsilon_ unset) */
  float target_density;        /* SkyPlaceDB targetDensity_ 1.0; must exceed the utilisation */
  float scaling_factor;        /* coordinates are divided by it on import and multiplied back on
                                  export (the reference's bookshelf dbu; 1 for DEF designs) */
  int32_t solver;              /* GP_SOLVER_*                                        */
  int32_t initial_place_only;  /* gp_place_design stops after the initial placement  */
  int32_t num_bin_x;           /* bin counts; 0 = createBins' rule (at most 1024 per side) */
  int32_t num_bin_y;
  int32_t precision;           /* DENS_PREC_* of the Poisson transforms (default BF16) */
  uint64_t seed;               /* of the initial placement's normal samples (default 1) */
  int32_t log_freq;            /* 50 */
  int32_t verbose;             /* 0: silent, 1: reference-style progress on stdout */
} GpParams;

/* Per-iteration trace (caller-allocated arrays of `cap` entries, any may be
 * NULL): values after iteration k's updateParameters. */
typedef struct GpTrace {
  int32_t cap;
  float* hpwl;       /* HPWL at the predicted point (computeHPWL)          */
  float* overflow;   /* density overflow at the predicted point           */
  float* lambda;     /* lambda after updateLambda                         */
  float* step;       /* accepted step length (Nesterov) / alpha (Adam)    */
  int32_t* backtracks; /* objective evaluations of the iteration (1: the
                          reference's repeated backtracking trials are equal) */
} GpTrace;

typedef struct GpStats {
  int32_t num_movable;     /* movable instances                            */
  int32_t num_fillers;     /* createFillers' count                          */
  int32_t num_bin_x, num_bin_y;
  float filler_width, filler_height;
  int32_t iterations;      /* the loop's final iter (NesterovSolver::solve) */
  int32_t converged;       /* checkConvergence() ended the loop             */
  int32_t evaluations;     /* objective + gradient evaluations              */
  float hpwl;              /* TargetFunction hpwl_ at the end               */
  float overflow;          /* overflow_ at the end                          */
  float lambda;
  float gamma_inv;
  float reference_hpwl;    /* in effect (after reference_hpwl_ratio)        */
  double hpwl_initial_db;  /* HPWL of the design's positions (pins as placed here) */
  double hpwl_db;          /* HPWL after gp_export's rounding               */
  double t_init_s;         /* gp_create + initial placement                */
  double t_solve_s;        /* gp_init_solver + loop                         */
} GpStats;

typedef struct GpHandle GpHandle;

const char* gp_last_error(void);
void gp_default_params(GpParams* p);

/* SkyPlaceDB::init: movable / fixed cells and pins from the design, bins,
 * fillers, density sizes and fixed overlap, the WA and density evaluators;
 * positions start at the design's (centres of inst_lx / inst_ly). */
int gp_create(const LgDesign* d, const GpParams* p, GpHandle** out);
/* InitialPlacer::doRandomInit: each movable cell's centre is drawn from
 * N(die centre, (0.05 die width)^2) x N(die centre, (0.05 die height)^2),
 * fillers mirrored about the centre (getMirrorX / getMirrorY), then moved
 * inside the die.  Deterministic in params.seed: sample k of a cell comes
 * from splitmix64(seed + 2 * index + k) by the Box-Muller transform. */
int gp_initial_place(GpHandle* h);
/* NesterovSolver::initSolver (getInitialGrad, moveBackward, Lipschitz step,
 * updateParameters) or AdamSolver::initSolver. */
int gp_init_solver(GpHandle* h);
/* Up to n iterations of the loop body; *done = 1 when checkConvergence()
 * ended the loop or max_opt_iter was reached. */
int gp_iterate(GpHandle* h, int32_t n, int32_t* done);
/* gp_init_solver + gp_iterate to the end (NesterovSolver::solve). */
int gp_solve(GpHandle* h, GpTrace* trace);
/* Host copies of the current centres (cur_var) of every movable cell and
 * filler ([num_movable + num_fillers] each); either may be NULL. */
int gp_get_positions(GpHandle* h, float* cx, float* cy);
/* Overwrite the current centres (both the solution and the predicted point). */
int gp_set_positions(GpHandle* h, const float* cx, const float* cy);
/* The solver's state for step-by-step checks: scalars[0..7] = lambda,
 * gamma_inv, step length, curA (Nesterov's a_k), hpwl, prev hpwl, overflow,
 * iteration; each array may be NULL, else [2 * (num_movable + num_fillers)]
 * (x then y): cur_var, cur_pre_var, cur_pre_grad. */
int gp_get_state(GpHandle* h, float* scalars, float* cur_var, float* cur_pre_var, float* cur_pre_grad);
/* SkyPlaceDB::exportDB: lower-left corners of the movable instances rounded
 * to dbu, written at their instance index into out_lx / out_ly
 * ([num_insts]; fixed instances keep the design's position). */
int gp_export(GpHandle* h, int32_t* out_lx, int32_t* out_ly);
int gp_get_stats(GpHandle* h, GpStats* s);
int gp_destroy(GpHandle* h);

/* SkyPlace::run on a design: gp_create, gp_initial_place, gp_solve (unless
 * initial_place_only), gp_export into out_lx / out_ly ([num_insts]). */
int gp_place_design(const LgDesign* d, const GpParams* p, int32_t* out_lx, int32_t* out_ly, GpStats* stats,
                    GpTrace* trace);

#ifdef __cplusplus
}
#endif

#endif /* SKYLINE_GP_H */
