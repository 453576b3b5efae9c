// FP64 instantiation of the step kernels, with IEEE division throughout: the FP64 build is
// the GPU twin of the FP64 oracle and follows its rounding.
#include "ibis_step_kernels.hpp"

namespace ibis {

int launch_f64(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
               int64_t slot_pitch, const double* table, const IbisParams& P,
               const StepConst& K, int64_t ncell, int copy_in, hipStream_t st) {
  return launch<double, double, double>(in, out, ws, slot, slot_pitch, table, P, K, ncell, copy_in, st);
}

int launch_ext_f64(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
                   int64_t slot_pitch, const double* tables, const IbisParams& P,
                   const StepConst& K, const StepExt& E, int64_t ncell, int copy_in,
                   hipStream_t st) {
  return launch_ext<double, double, double>(in, out, ws, slot, slot_pitch, tables, P, K, E,
                                            ncell, copy_in, st);
}

}  // namespace ibis
