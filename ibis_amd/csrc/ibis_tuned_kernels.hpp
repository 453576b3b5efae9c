// Kernels of the parameter-estimation step (include/ibis_adjoint.h): one thread per
// (pair, direction), the per-thread bodies of ibis_tuned_cell.hpp.  Instantiated per precision
// in ibis_tuned_fast.hip (mixed, reciprocal division like the product step) and
// ibis_tuned_f64.hip (FP64, IEEE division), each for plain values and for dual numbers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdio.h>

#include "ibis_error.hpp"
#include "ibis_step_kernels.hpp"
#include "ibis_tuned_cell.hpp"

namespace ibis {

#define TUNED_THREAD(g, th)                                              \
  const int64_t t_ = (int64_t)blockIdx.x * kBlock + threadIdx.x;         \
  if (t_ >= (g).nthread) return;                                         \
  const TunedThread th = tuned_thread_of((g), t_)

template <class S, class R, class D, class Q>
__global__ __launch_bounds__(kBlock, kWavesSurface) void tuned_surface_kernel(
    S* a, void* ws, TunedArgs g, IbisParams base, StepConst K) {
  TUNED_THREAD(g, th);
  tuned_surface<S, R, D, Q>(a, ws, g, th, base, K);
}

template <class S, class R, class D, class Q>
__global__ __launch_bounds__(kBlock, kWavesCanopy) void tuned_canopy_kernel(
    S* a, void* ws, TunedArgs g, IbisParams base, StepConst K) {
  TUNED_THREAD(g, th);
  tuned_canopy<S, R, D, Q>(a, ws, g, th, base, K);
}

template <class S, class R, class D, class Q>
__global__ __launch_bounds__(kBlock, kWavesGround) void tuned_ground_kernel(
    S* a, void* ws, TunedArgs g, IbisParams base, StepConst K) {
  TUNED_THREAD(g, th);
  tuned_ground<S, R, D, Q>(a, ws, g, th, base, K);
}

template <class S>
__global__ __launch_bounds__(kBlock) void tuned_misfit_kernel(const S* a, TunedArgs g,
                                                              MisfitArgs m, const float* target,
                                                              float* flux, double* loss) {
  TUNED_THREAD(g, th);
  tuned_misfit<S>(a, g, th, m, target, flux, loss);
}

template <class S, class R, class D, class Q>
inline int tuned_launch(S* a, void* ws, const TunedArgs& g, const IbisParams& base,
                        const StepConst& K, const MisfitArgs& m, const float* target,
                        float* flux, double* loss, hipStream_t st) {
  const dim3 grid((unsigned)((g.nthread + kBlock - 1) / kBlock)), block(kBlock);
  hipLaunchKernelGGL((tuned_surface_kernel<S, R, D, Q>), grid, block, 0, st, a, ws, g, base, K);
  hipLaunchKernelGGL((tuned_canopy_kernel<S, R, D, Q>), grid, block, 0, st, a, ws, g, base, K);
  hipLaunchKernelGGL((tuned_ground_kernel<S, R, D, Q>), grid, block, 0, st, a, ws, g, base, K);
  if (flux || (target && loss))
    hipLaunchKernelGGL((tuned_misfit_kernel<S>), grid, block, 0, st, (const S*)a, g, m, target,
                       flux, loss);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(err_buf(), 512, "ibis_tuned_hour: launch failed: %s", hipGetErrorString(e));
    return IBIS_E_LAUNCH;
  }
  return IBIS_OK;
}

}  // namespace ibis
