// FP64 instantiations of the parameter-estimation kernels, in values and in dual numbers,
// with IEEE division: the reference for finite-difference checks of the mixed build.
#include "ibis_tuned_kernels.hpp"

namespace ibis {

int tuned_launch_f64(void* a, void* ws, const TunedArgs& g, const IbisParams& base,
                     const StepConst& K, const MisfitArgs& m, const float* target, float* flux,
                     double* loss, bool tangent, hipStream_t st) {
  if (tangent)
    return tuned_launch<Dual<double>, Dual<double>, Dual<double>, Dual<double>>(
        (Dual<double>*)a, ws, g, base, K, m, target, flux, loss, st);
  return tuned_launch<double, double, double, double>((double*)a, ws, g, base, K, m, target,
                                                      flux, loss, st);
}

}  // namespace ibis
