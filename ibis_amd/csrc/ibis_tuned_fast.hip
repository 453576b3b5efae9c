// Mixed-precision instantiations of the parameter-estimation kernels (FP32 state and bulk
// arithmetic, FP64 solves), in values and in dual numbers.  Built with the product step's
// reciprocal division, so the value run performs the product step's arithmetic.
#include "ibis_tuned_kernels.hpp"

namespace ibis {

int tuned_launch_mixed(void* a, void* ws, const TunedArgs& g, const IbisParams& base,
                       const StepConst& K, const MisfitArgs& m, const float* target, float* flux,
                       double* loss, bool tangent, hipStream_t st) {
  if (tangent)
    return tuned_launch<Dual<float>, Dual<float>, Dual<double>, Dual<double>>(
        (Dual<float>*)a, ws, g, base, K, m, target, flux, loss, st);
  return tuned_launch<float, float, double, double>((float*)a, ws, g, base, K, m, target, flux,
                                                    loss, st);
}

}  // namespace ibis
