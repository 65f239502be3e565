// Internal entry points shared by the translation units of the legalizer.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/skyline_lg.h"

// QP refine of CudaPostProcess (CudaPostProcess.cu:36-66) on the handle's
// device state; implemented in lg_qp.hip.
int lg_qp_refine(LgHandle* h, hipStream_t s);
