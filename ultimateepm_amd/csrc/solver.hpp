// solver.hpp — launch interface between the host C ABI (uepm_abi.cpp) and the
// device stages (solver_f64.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "epm_physics.hpp"

namespace uepm {

hipError_t launch_assemble(const DevParams& P, const int32_t* dG, const double* dk, int64_t nk, double2* dH, int ld,
                           hipStream_t s);
hipError_t launch_tridiag(double2* dH, int n, int ld, int64_t nk, double* dd, double* de, double2* dtau,
                          hipStream_t s);
hipError_t launch_tri_eig(const double* dd, const double* de, int n, int nb, int64_t nk, double* devals, double* dZ,
                          double* dW, int want_vectors, hipStream_t s);
hipError_t launch_backtransform(const double2* dH, int n, int ld, const double2* dtau, const double* dd,
                                const double* de, const double* devals, const double* dZ, double2* dX, double* dgrads,
                                const DevParams& P, const double* dk, const int32_t* dG, int nb, int64_t nk,
                                hipStream_t s);

}  // namespace uepm
