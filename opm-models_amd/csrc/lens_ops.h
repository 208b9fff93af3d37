// lens_ops.h -- host launchers of the two-phase immiscible element path
// (lens_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "lens_physics.h"

namespace omodels {

void launchLensStorage(const LensParams& P, const double* pv, const uint8_t* inLens, double* st, hipStream_t s);
void launchLensAssemble(const LensParams& P, const double* pv, const uint8_t* inLens, const double* st1, double dt,
                        double* res, double* jac, const int* diagOff, const int* nbOff, hipStream_t s);
void launchLensUpdate(long m, double* pv, const double* dx, int* bad, hipStream_t s);

}  // namespace omodels
