/*
 * skyline_lg.h -- C ABI of the MI355X (gfx950) ADMM negotiation legalizer.
 *
 * Plain C: fixed-width integers, raw pointers and sizes, no torch or HIP types
 * in any signature (a stream is passed as `void*`, a hipStream_t value or NULL
 * for the library's own stream).  Every entry point returns an int status
 * (LG_OK == 0); on failure lg_last_error() returns a message for the calling
 * thread.  The library never calls exit() or abort().
 *
 * Two layers, each replacing one layer of the reference:
 *
 *  1. Design level -- lg_run_design() replaces legalizer::Legalizer::run()
 *     (reference: src/skyline/legalizer/src/Legalizer.cpp:197-228): import of
 *     cells/regions (DatabaseIO.cpp:31-379), pixel grid (Grid.cpp:24-407),
 *     preplace (Preplace.cpp:18-207), then the device layer below, then the
 *     statistics (Legalizer.cpp:439-476) and checkPlace (CudaDatabase.cu:326-353).
 *
 *  2. Device level -- lg_create / lg_solve / lg_postprocess / lg_check /
 *     lg_copy_out / lg_destroy replace the CUDA layer that Legalizer drives:
 *     CudaDatabase (database/CudaDatabase.h:16-49), CudaAdmmLegalizer
 *     (admm_legalizer/CudaAdmmLegalizer.h:19-33) and CudaPostProcess
 *     (post_process/CudaPostProcess.h:12-20).
 *
 * Units: "dbu" are design database units; "grid" coordinates count sites (x)
 * and rows (y) from the core's lower-left corner.  Pixel arrays are x-major:
 * pixel_id = x * y_size + y (reference AdmmKernels.hpp:74).
 */
#ifndef SKYLINE_LG_H
#define SKYLINE_LG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LG_OK = 0,
  LG_ERR_ARG = 1,        /* bad argument / inconsistent sizes            */
  LG_ERR_DEVICE = 2,     /* HIP runtime error, or no usable gfx950 device */
  LG_ERR_DESIGN = 3,     /* design cannot be legalized (movable macro, no space) */
  LG_ERR_MEMORY = 4,     /* allocation failed or over the memory budget   */
  LG_ERR_STATE = 5       /* call out of order (e.g. copy_out before solve) */
};

/* Row orientations (DEF).  N and FN rows have VDD on top of row y
 * (reference Grid.cpp:337-341: is_vdd_up = 1), all others VSS on top. */
enum { LG_ORIENT_N = 0, LG_ORIENT_S = 1, LG_ORIENT_W = 2, LG_ORIENT_E = 3,
       LG_ORIENT_FN = 4, LG_ORIENT_FS = 5, LG_ORIENT_FW = 6, LG_ORIENT_FE = 7 };

/* Sorting policy / algorithm (reference LGHyperParameters.h:4-5). */
enum { LG_SORT_POLICY_RANDOM = 0, LG_SORT_POLICY_SIZE = 1, LG_SORT_POLICY_OVERFLOW = 2 };
enum { LG_SORT_S_MERGE = 0, LG_SORT_M_MERGE = 1, LG_SORT_S_BITONIC = 2,
       LG_SORT_M_BITONIC = 3, LG_SORT_CUB = 4 };

/* Hyper-parameters; field for field the reference's LGHyperParameters
 * (legalizer/src/LGHyperParameters.h:7-52).  lg_default_params() fills the
 * defaults of Legalizer::initParameters (Legalizer.cpp:47-93). */
typedef struct LgParams {
  int32_t flag_iccad17;
  int32_t standard_admm;
  int32_t log_freq;
  int32_t num_cpu_threads;
  int32_t max_disp_in_row_height;
  int32_t max_admm_iter;
  int32_t admm_terminate_thr;
  int32_t x_hint;
  int32_t y_hint;
  int32_t iter_update_partition;
  float   max_disp_coeff;
  float   tech_penalty;
  float   edge_spacing_penalty;
  int32_t iter_threshold_dual_update;
  int32_t iter_update_dual_step;
  int32_t iter_perturbation_on;
  float   coeff_to_init_rho;
  float   coeff_to_init_dual;
  float   coeff_to_update_dual_step;
  int32_t qp_admm_on;
  int32_t qp_admm_dp_site_align;
  int32_t qp_admm_max_iter;
  int32_t qp_admm_min_iter;
  float   qp_admm_rho;
  int32_t use_adaptive_search;
  int32_t stagnation_window;  /* 1..64 with use_adaptive_search (else LG_ERR_ARG) */
  int32_t min_iter_to_call_adaptive_search;
  int32_t cooltime_to_call_adaptive_search;
  float   stagnation_threshold;
  float   ovf_ratio_to_call_adaptive_search;
  float   coeff_to_neglect_displace;
  int32_t random_sorting_seed;
  int32_t sorting_policy;     /* LG_SORT_POLICY_* */
  int32_t sorting_algorithm;  /* LG_SORT_*        */
  int32_t verbose;            /* 0: silent, 1: reference-style progress on stdout */
} LgParams;

/* In-memory design: the subset of the reference's db::dbDatabase that the
 * legalizer reads (core box, rows, instances, fence groups).  All in dbu. */
typedef struct LgDesign {
  int32_t dbu_per_micron;
  int32_t core_lx, core_ly, core_ux, core_uy;
  int32_t num_rows;
  const int32_t* row_lx;          /* [num_rows] */
  const int32_t* row_ly;
  const int32_t* row_dx;          /* row width  */
  const int32_t* row_dy;          /* row height */
  const int32_t* row_site_width;
  const int32_t* row_orient;      /* LG_ORIENT_* */
  int32_t num_insts;
  const int32_t* inst_lx;         /* [num_insts] global-placement position */
  const int32_t* inst_ly;
  const int32_t* inst_dx;
  const int32_t* inst_dy;
  const uint8_t* inst_fixed;      /* 1: FIXED/PLACED-fixed instance          */
  const uint8_t* inst_macro;      /* 1: master of class BLOCK                */
  const int32_t* inst_group;      /* -1 or index of the fence group          */
  const uint8_t* inst_ground_at_bottom; /* master's VSS pin is below its VDD pin */
  int32_t num_groups;
  int32_t num_region_boxes;       /* fence boxes, any number per group       */
  const int32_t* box_group;
  const int32_t* box_lx;
  const int32_t* box_ly;
  const int32_t* box_ux;
  const int32_t* box_uy;
  /* Optional netlist (num_nets = 0: none) for the HPWL before and after
   * legalization (Legalizer::computeHpwlFromDB, Legalizer.cpp:402-437).  The
   * pins of net k are net_start[k] .. net_start[k+1]-1; pin p sits at
   * (pin_dx, pin_dy) from the lower-left corner of instance pin_inst[p] in
   * orientation N (the master pin's average position, dbITerm::getAvgXY), or,
   * with pin_inst[p] = -1, is an IO pin at the absolute position (pin_dx,
   * pin_dy).  A cell the legalizer flips (FS) mirrors its pins' y. */
  int32_t num_nets;
  const int32_t* net_start;       /* [num_nets + 1] */
  const int32_t* pin_inst;
  const int32_t* pin_dx;
  const int32_t* pin_dy;
  /* Optional inputs of Legalizer::setSizeByFile / setICCAD17Constraint
   * (Legalizer.h:74-75), NULL for none: instance names [num_insts]; a size
   * file of lines "name=width_in_sites=height_in_rows" that resizes the named
   * movable cells (DatabaseIO.cpp:88-131, Util.cpp:220-234); an ICCAD 2017
   * placement.constraints file (maximum_utilization=U%,
   * maximum_movement=Mrows; Util.cpp:190-217), which turns on the contest
   * metric in the statistics. */
  const char* const* inst_name;
  const char* size_file;
  const char* constraint_file;
  /* Optional ICCAD 2017 technology (num_masters = 0: none): what
   * Legalizer::importTechInfo reads from the db (DatabaseIO.cpp:381-572).
   * inst_master: the LEF master of each instance (-1: none); per master its
   * left / right edge type (LEF58_EDGETYPE) and signal pins, each pin the
   * routing-layer index of its ports and the bounding box of their shapes
   * (dbu from the master's origin; the POWER / GROUND pins are left out);
   * the special-net wire segments (routing-layer index, end points, width,
   * dbu); the cell edge spacing table (LEF58_CELLEDGESPACINGTABLE: edge types
   * and spacing in microns).  The routing-layer index counts ROUTING layers
   * in LEF order from 0. */
  int32_t num_masters;
  const int32_t* inst_master;       /* [num_insts]                     */
  const int32_t* master_ledge;      /* [num_masters]                   */
  const int32_t* master_redge;
  const uint8_t* master_is_block;   /* class BLOCK: no tech info (DatabaseIO.cpp:455) */
  const int32_t* master_pin_start;  /* [num_masters + 1]               */
  const int32_t* mpin_layer;        /* [pins]                          */
  const int32_t* mpin_lx;
  const int32_t* mpin_ly;
  const int32_t* mpin_ux;
  const int32_t* mpin_uy;
  int32_t num_pwires;
  const int32_t* pwire_layer;       /* [num_pwires]                    */
  const int32_t* pwire_x1;
  const int32_t* pwire_y1;
  const int32_t* pwire_x2;
  const int32_t* pwire_y2;
  const int32_t* pwire_width;
  int32_t num_edge_spacings;
  const int32_t* es_type1;          /* [num_edge_spacings]             */
  const int32_t* es_type2;
  const double* es_spacing_um;
} LgDesign;

/* Violation counts of CudaDatabase::checkPlace (objects/LGVioStat.h), then
 * the strict legality counts, which have no counterpart in the reference's
 * checkPlace: they make the placement filters of computePenaltySum
 * (AdmmKernels.hpp:355-373) checkable after the fact.  Each strict count is
 * a number of movable cells. */
typedef struct LgVioStat {
  int32_t total_ovf;            /* sum_p max(usage-1, 0): overlap             */
  int32_t num_row_orient_vio;   /* countRowOrientVioKernel                    */
  int32_t num_pin_short_vio;
  int32_t num_pin_access_vio;
  int32_t num_edge_spacing_vio;
  int32_t num_site_vio;         /* footprint covers a pixel that is no valid
                                   site (blockage, macro, outside every row) */
  int32_t num_fence_vio;        /* footprint covers a valid site of another
                                   fence group (or of a fence, for a cell of
                                   the default region)                        */
  int32_t num_core_vio;         /* footprint leaves the core (the pixel grid) */
  int32_t num_rail_vio;         /* even-height cell whose bottom row has the
                                   wrong power rail (not flippable)           */
} LgVioStat;

typedef struct LgStats {
  int32_t iterations;         /* ADMM iterations executed                          */
  int32_t converged;          /* 1 if overflow reached admm_terminate_thr          */
  int64_t final_overflow;     /* sum_p max(usage-1,0) after the last iteration     */
  int64_t total_disp_dbu;     /* sum over movable cells |dx|+|dy| (exact)          */
  int64_t max_disp_dbu;
  double  total_disp_um;      /* Legalizer::computeStat (Legalizer.cpp:439-476)    */
  double  avg_disp_um;        /*   averaged over ALL design instances, as there    */
  double  max_disp_um;
  float   rho;                /* initial penalty rho                               */
  float   initial_dual;       /* initial lambda                                    */
  int32_t num_cells;          /* movable standard cells legalized                  */
  int32_t num_adaptive_calls; /* adaptive-search primal updates executed           */
  LgVioStat vio;
  double  t_import_s;         /* importDB + grid                                   */
  double  t_preplace_s;       /* initialLegalize                                   */
  double  t_admm_s;           /* device upload + ADMM (runAdmmLegalize)            */
  double  t_post_s;           /* runPostProcess                                    */
  double  t_total_s;          /* whole lg_run_design                               */
  /* lg_run_design with a netlist: HPWL of the global placement and of the
   * result (dbu / dbu_per_micron) and hpwl_after / hpwl_before * 100 - 100
   * (Util.cpp:113); zeros without nets or at the device level */
  double  hpwl_before_um;
  double  hpwl_after_um;
  double  delta_hpwl_pct;
  /* lg_run_design with a constraint file: its limits and the ICCAD 2017
   * contest metric (Legalizer::printICCAD17Metric, Util.cpp:141-188):
   * S_am = sum over heights h of (sum of displacements in sites of the
   * h-row cells / their count) / #heights / (row height / site width),
   * M_max = max displacement / row height, S = (1 + dHPWL/100 + (N_p + N_e) /
   * #cells) * (1 + M_max / 100) * S_am */
  double  max_util_constraint;   /* percent */
  double  max_disp_constraint;   /* dbu */
  double  iccad17_s_am;
  double  iccad17_m_max;
  double  iccad17_score;
} LgStats;

/* Outputs of lg_run_design, all caller-allocated.  Per-instance arrays have
 * num_insts entries; fixed instances are copied through unchanged. */
typedef struct LgResult {
  int32_t* out_lx;            /* legalized lower-left x (dbu)                    */
  int32_t* out_ly;
  uint8_t* out_flipped;       /* 1 if the cell gets orientation FS (rotation)    */
  int32_t* ovf_trace;         /* optional [ovf_trace_cap]: overflow per iteration */
  int32_t  ovf_trace_cap;
  LgStats  stats;
} LgResult;

/* ---- device-level inputs (what CudaDatabase copies to the GPU) ---------- */

typedef struct LgGridDesc {
  int32_t x_size, y_size;          /* sites x rows                                 */
  int32_t site_width, row_height;  /* dbu                                          */
  int32_t core_lx, core_ly;        /* dbu                                          */
  const int32_t* pixel_valid;      /* [x_size*y_size] 0/1                          */
  const int32_t* pixel_group;      /* [x_size*y_size] -1 default, -2 blockage, >=0 fence */
  const int32_t* pixel_region;     /* [x_size*y_size] region id or -1              */
  const int32_t* row_vdd_up;       /* [y_size] 0/1                                 */
  int32_t num_regions;             /* shrunk region bounds (adaptive search)       */
  const int32_t* region_x_min;
  const int32_t* region_y_min;
  const int32_t* region_x_max;
  const int32_t* region_y_max;
  /* Optional ICCAD 2017 technology (NULL / 0: none).  pixel_power_layer
   * [x_size*y_size], x-major: routing-layer index of the power metal over
   * the pixel, or -1 (Grid::markPowerMetalLayer, Grid.cpp:345-362, from the
   * special nets on layers 1 and 2, Legalizer::importPowerMetals,
   * DatabaseIO.cpp:528-572).  edge_spacing_rules[num_edge_rules]: the
   * reference's 1-D rule table, spacing in sites between edge types a and b
   * at index a + b (Legalizer::importEdgeSpcingRules, DatabaseIO.cpp:416-441);
   * types are 0 (none), 1 or 2, so indices up to 4 are read. */
  const int32_t* pixel_power_layer;
  int32_t num_edge_rules;
  const int32_t* edge_spacing_rules;
} LgGridDesc;

typedef struct LgCellsSoA {
  int32_t num_cells;
  const int32_t* orig_lx_dbu;      /* global-placement position                    */
  const int32_t* orig_ly_dbu;
  const int32_t* lx_dbu;           /* preplaced, on the site/row grid              */
  const int32_t* ly_dbu;
  const int32_t* w_grid;           /* width in sites                               */
  const int32_t* h_grid;           /* height in rows                               */
  const int32_t* group;            /* -1 or fence group                            */
  const int32_t* ground_at_bottom; /* 0/1                                          */
  /* Optional cell technology (NULL / 0: none), what CudaDatabase copies from
   * LGCellTechInfo (CudaDatabase.cu:85-148; DatabaseIO.cpp:450-512): the
   * left / right edge type of each cell's master (0, 1 or 2), the master
   * index of each cell and, per master, its signal-pin boxes in grid units,
   * each corner flattened x * h_grid + y inside the cell, with the pin's
   * routing-layer index.  With any of these the candidate cost gains the
   * pin-access and edge-spacing terms (AdmmKernels.hpp:193-331, 806-830). */
  const int32_t* ledge_type;       /* [num_cells] */
  const int32_t* redge_type;       /* [num_cells] */
  const int32_t* macro_id;         /* [num_cells], index < num_macros             */
  int32_t num_macros;
  const int32_t* macro_pin_start;  /* [num_macros + 1]                            */
  const int32_t* pin_bgn;          /* [macro_pin_start[num_macros]]               */
  const int32_t* pin_end;
  const int32_t* pin_layer;
} LgCellsSoA;

typedef struct LgHandle LgHandle;

/* Fill `p` with the reference defaults (Legalizer.cpp:47-93). */
void lg_default_params(LgParams* p);

/* Message of the last failed call on this thread ("" if none). */
const char* lg_last_error(void);

/* Whole legalization of `d` (Legalizer::run). */
int lg_run_design(const LgDesign* d, const LgParams* p, LgResult* r);

/* Import + preplace only (no device work), for parity checks of the host
 * side.  shape[4] receives x_size, y_size, num_cells, num_regions; call once
 * with valid == NULL to query them.  Then: pixel maps (x-major, x_size*y_size),
 * vdd_up[y_size], region_bounds[4*num_regions] (shrunk grid bounds x0 y0 x1 y1)
 * and the preplaced movable-cell positions (dbu, movable-cell order).
 * Replaces importDB + initialLegalize (DatabaseIO.cpp:31-77, Preplace.cpp:18-72). */
int lg_import_and_preplace(const LgDesign* d, const LgParams* p, int32_t* shape, int32_t* valid,
                           int32_t* group, int32_t* region, int32_t* vdd_up, int32_t* region_bounds,
                           int32_t* cell_lx, int32_t* cell_ly);
/* The same import with the pixel grid built and the cells preplaced on the
 * device: the path lg_run_design takes when the adaptive search is off and
 * no two fence boxes overlap (else LG_ERR_STATE).  shape[4] as above (a
 * query with valid == NULL works for every design); valid / group are the
 * x-major maps, group -1 where a pixel is invalid (the device keeps no group
 * there); cell_lx / cell_ly the preplaced positions.  No counterpart in the
 * reference (whose grid and preplace are host code). */
int lg_import_and_preplace_device(const LgDesign* d, const LgParams* p, int32_t* shape, int32_t* valid,
                                  int32_t* group, int32_t* cell_lx, int32_t* cell_ly);

/* Device layer.  lg_create uploads the grid and the preplaced cells
 * (CudaDatabase ctor + copyFromHostCells, CudaDatabase.cu:24-236). */
int lg_create(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, LgHandle** out);
/* ADMM negotiation loop incl. final refine (CudaAdmmLegalizer::solve,
 * CudaAdmmLegalizer.cu:250-319). */
int lg_solve(LgHandle* h, void* stream);
/* Rotation and, if qp_admm_on, the QP refine (CudaPostProcess::doPostProcess,
 * CudaPostProcess.cu:27-66). */
int lg_postprocess(LgHandle* h, void* stream);
/* checkPlace (CudaDatabase.cu:326-353). */
int lg_check(LgHandle* h, LgVioStat* out);
/* Grid positions and rotation flags (CudaDatabase::copyToHostCells,
 * CudaDatabase.cu:274-309); any pointer may be NULL. */
int lg_copy_out(LgHandle* h, int32_t* lx_grid, int32_t* ly_grid, int32_t* need_rotation);
/* Solve statistics (iterations, overflow, rho, lambda0, displacement). */
int lg_get_stats(LgHandle* h, LgStats* out);
/* Pixel state after the last call: usage and dual, x-major; any may be NULL. */
int lg_get_pixel_state(LgHandle* h, int32_t* usage, float* dual);
/* Overflow per executed iteration, up to cap entries; returns count in *n. */
int lg_get_ovf_trace(LgHandle* h, int32_t* trace, int32_t cap, int32_t* n);
int lg_destroy(LgHandle* h);

/* Instrumentation; no counterpart in the reference.  With enable != 0 every
 * later lg_solve brackets each ADMM iteration's partition-kernel launches
 * (small LDS class, large LDS class) with HIP events on the solve's stream;
 * lg_profile_read returns the summed milliseconds and launch counts per class
 * since the last enable.  Adds three event records per iteration. */
int lg_profile_partition(LgHandle* h, int enable);
int lg_profile_read(LgHandle* h, double* ms /*[2]*/, int64_t* launches /*[2]*/);
/* HIP graphs for the ADMM iterations of lg_solve; no counterpart in the
 * reference.  Off by default (kernel launches).  When enabled, the first
 * solve captures eight iterations (every kernel of the iteration, with the
 * dual pass on its side branch) into one graph, which each later batch of
 * iterations replays; profiled solves (lg_profile_partition) launch
 * eagerly.  Results are the same either way. */
int lg_use_graphs(LgHandle* h, int32_t enable);
/* Per-phase clock ticks of a build compiled with LG_PHASE_TIMERS (zeros
 * otherwise); no counterpart in the reference. */
int lg_debug_phase_ticks(LgHandle* h, uint64_t* ticks, int n);

/* ---- request server (no counterpart in the reference) ------------------
 *
 * A continuous-batching legalization service on one GPU.  Requests (whole
 * designs, ECO windows, sub-designs: the device-level inputs of lg_create)
 * are submitted from any thread and queued in arrival order.  A server
 * thread admits the head of the queue whenever a slot is free and its device
 * buffers fit the memory budget, keeps each admitted request's cell arrays,
 * pixel grid and dual prices resident until it converges, and runs one ADMM
 * iteration of every resident request per batched kernel launch, so requests
 * that arrive while others run share the device with them.  Each request
 * computes exactly what lg_create + lg_solve + lg_postprocess compute for it
 * alone (same kernels, same order of operations, per-request rho / lambda /
 * iteration control), so its result does not depend on what it is batched
 * with.  Finished requests are retired between batches (final refine,
 * rotation, optional QP refine, checkPlace, positions copied to the host). */

enum { LG_SRV_QUEUED = 0, LG_SRV_RUNNING = 1, LG_SRV_DONE = 2, LG_SRV_FAILED = 3 };

typedef struct LgSrvConfig {
  int32_t max_slots;          /* requests resident on the device at once (1..1024)   */
  int32_t iters_per_sync;     /* ADMM iterations launched between host checks (1..1024) */
  int64_t mem_budget_bytes;   /* device memory the server may hold for requests     */
  int64_t pool_chunk_bytes;   /* granularity in which that memory is taken          */
  int32_t use_graphs;         /* 1: each ADMM iteration of the resident requests is
                                 one HIP graph launch, recaptured when the set of
                                 resident requests changes; 0 (default): kernel
                                 launches.  Results are the same either way.     */
  int32_t short_cells;        /* > 0: requests with fewer cells iterate in a second
                                 lane with its own streams and batch cadence, so
                                 they are not paced by the long requests' launches;
                                 max_slots bounds the resident requests of both
                                 lanes together.  0: one lane.                    */
  int32_t max_overtakes;      /* admission order.  0: arrival order.  > 0: the
                                 smallest queued request that fits goes first, but
                                 no request is passed over by more than this many
                                 later requests: once the oldest queued request has
                                 been, requests are admitted in arrival order until
                                 it is in.  Results do not depend on either field. */
  int32_t reserved;           /* 0 */
} LgSrvConfig;

typedef struct LgSrvResult {
  int32_t* out_lx;            /* [num_cells] grid x (sites), or NULL                */
  int32_t* out_ly;            /* [num_cells] grid y (rows), or NULL                 */
  int32_t* out_rot;           /* [num_cells] need_rotation, or NULL                 */
  int32_t  num_cells;         /* cells of the request (written)                     */
  int32_t  status;            /* LG_OK or the request's error code (written)        */
  LgStats  stats;             /* as lg_get_stats after lg_check (written)           */
  double   queued_s;          /* submit -> admitted to a slot                       */
  double   run_s;             /* admitted -> result on the host                     */
  double   latency_s;         /* submit -> result on the host                       */
} LgSrvResult;

typedef struct LgSrvCounters {
  int64_t submitted, completed, failed;
  int64_t batches;            /* batches ended (host checks), summed over lanes     */
  int64_t iterations;         /* batched ADMM iterations launched, over lanes       */
  int32_t peak_slots;         /* most requests resident at once                     */
  int32_t active, queued;
  int64_t pool_reserved_bytes, pool_peak_bytes;
  int64_t graph_captures;     /* iteration graphs captured (use_graphs)             */
} LgSrvCounters;

typedef struct LgServer LgServer;

/* max_slots 128, iters_per_sync 16, budget 200 GB, chunks of 256 MiB, no graphs,
 * two lanes split at 20000 cells, smallest first with max_overtakes 64.
 * max_slots trades capacity for tail latency: each resident request advances
 * one iteration per batched iteration of its lane, whose duration grows with
 * the lane's resident work, while the capacity grows with the residency. */
void lg_srv_default_config(LgSrvConfig* c);
/* Starts the server thread on the calling thread's current device. */
int lg_srv_create(const LgSrvConfig* c, LgServer** out);
/* Validates and prepares the request on the calling thread (the host work of
 * lg_create), queues it and returns its ticket; the inputs may be freed on
 * return.  LG_ERR_MEMORY if the request alone exceeds the budget. */
int lg_srv_submit(LgServer* s, const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, int64_t* ticket);
/* LG_SRV_* state of a ticket that has not been collected by lg_srv_wait. */
int lg_srv_poll(LgServer* s, int64_t ticket, int32_t* state);
/* Blocks until the request is done, fills `r` and releases the ticket.
 * Returns the request's status (its message through lg_last_error).  One
 * thread per ticket: a second concurrent wait returns LG_ERR_STATE. */
int lg_srv_wait(LgServer* s, int64_t ticket, LgSrvResult* r);
int lg_srv_counters(LgServer* s, LgSrvCounters* out);
/* Instrumentation: with enable != 0, every later batched iteration brackets
 * its partition-kernel launches (the primal step of the lane's resident
 * requests) with HIP events on the lane's stream.  lg_srv_profile_read
 * returns, since the last enable: the summed milliseconds of those launches,
 * the number of bracketed iterations (over lanes), the algorithmic bytes of
 * the iterations the requests executed in them (12 per pixel + 24 per cell
 * and iteration) and, if ms_union is not NULL, the milliseconds during which
 * at least one lane's partition launches ran (the lanes' launches overlap,
 * so this is at most the sum). */
int lg_srv_profile(LgServer* s, int enable);
int lg_srv_profile_read(LgServer* s, double* ms, int64_t* iterations, double* bytes, double* ms_union);
/* Stops the server: queued and running requests are dropped; threads
 * blocked in lg_srv_wait return LG_ERR_STATE (or their result, for a request
 * that finished first) and lg_srv_destroy returns only after every such
 * thread has left lg_srv_wait.  Device memory is released.  Calls on the
 * server that start after lg_srv_destroy began are undefined. */
int lg_srv_destroy(LgServer* s);

#ifdef __cplusplus
}
#endif

#endif /* SKYLINE_LG_H */
