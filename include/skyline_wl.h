/*
 * skyline_wl.h -- C ABI of the MI355X (gfx950) weighted-average wirelength
 * gradient and the target function's preconditioned total gradient.
 *
 * Replaces skyplace::WireLengthGradient (reference:
 * src/skyline/skyplace/src/wirelength_gradient/WireLengthGradient.h:13-139,
 * WireLengthGradient.cu:261-568) and the jacobianPrecondition kernel of
 * TargetFunction::updatePointAndGetGrad (problem_instance/TargetFunction.cu:88-141,
 * ProblemInstanceKernel.hpp:74-100).  All array arguments named d_* are
 * device pointers; `stream` is a hipStream_t value or NULL (the handle's own
 * stream).  Every entry point returns 0 (WL_OK) or an error code, with a
 * message for the calling thread from wl_last_error().
 *
 * Sign convention as in the reference: the "gradient" of a pin is
 * -d(WA)/dx, the descent direction that TargetFunction adds to lambda times
 * the density force.
 */
#ifndef SKYLINE_WL_H
#define SKYLINE_WL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { WL_OK = 0, WL_ERR_ARG = 1, WL_ERR_DEVICE = 2, WL_ERR_MEMORY = 4 };

/* The netlist as WireLengthGradient::initForCUDAKernel lays it out
 * (WireLengthGradient.cu:404-568): pins numbered net by net. */
typedef struct WlNetlist {
  int32_t num_cells;           /* movable cells                                      */
  int32_t num_nets;
  int32_t num_pins;
  const int32_t* net_start;    /* [num_nets + 1]; net n owns pins net_start[n] .. net_start[n+1]-1 */
  const int32_t* pin2cell;     /* [num_pins] movable cell of the pin, or -1 (fixed cell, IO pin) */
  const float* pin_offset_x;   /* [num_pins] pin centre minus cell centre            */
  const float* pin_offset_y;
  const float* pin_x;          /* [num_pins] initial pin centres (kept for pins of no movable cell) */
  const float* pin_y;
  const float* net_weight;     /* [num_nets], or NULL for weight 1                   */
} WlNetlist;

typedef struct WlHandle WlHandle;

const char* wl_last_error(void);
/* WireLengthGradient(db): uploads the netlist (host -> device, synchronous). */
int wl_create(const WlNetlist* n, WlHandle** out);
/* updatePinCoordinates (WireLengthGradient.cu:368-381). */
int wl_update_pin_coordinates(WlHandle* h, const float* d_cell_cx, const float* d_cell_cy, void* stream);
/* computeGrad (WireLengthGradient.cu:275-366) at the stored pin centres;
 * d_grad_x / d_grad_y have num_cells entries and are overwritten. */
int wl_compute_grad(WlHandle* h, float gamma_inv, float* d_grad_x, float* d_grad_y, void* stream);
/* updatePinCoordinates followed by computeGrad (the pair TargetFunction
 * calls back to back), fused into one pass over the pins. */
int wl_compute_grad_at(WlHandle* h, const float* d_cell_cx, const float* d_cell_cy, float gamma_inv,
                       float* d_grad_x, float* d_grad_y, void* stream);
/* computeHPWL (WireLengthGradient.cu:383-401), summed in FP64; synchronises. */
int wl_compute_hpwl(WlHandle* h, double* hpwl, void* stream);
/* computeHPWL into the device double *d_out, enqueued on the stream without
 * waiting (no counterpart in the reference, which returns it to the host). */
int wl_compute_hpwl_async(WlHandle* h, double* d_out, void* stream);
/* Host copy of the stored pin centres (synchronises); either may be NULL. */
int wl_get_pins(WlHandle* h, float* pin_x, float* pin_y);
int wl_destroy(WlHandle* h);

/* jacobianPrecondition (ProblemInstanceKernel.hpp:74-100): for each of n
 * cells, g = (wl + lambda * dens) / max(int(num_pins) + lambda * precond, min_precond). */
int tf_precondition(int32_t n, float lambda, float min_precond, const float* d_num_pins,
                    const float* d_dens_precond, const float* d_wl_gx, const float* d_wl_gy,
                    const float* d_dens_gx, const float* d_dens_gy, float* d_gx, float* d_gy, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SKYLINE_WL_H */
