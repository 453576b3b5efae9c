// Mixed-precision and FP32 instantiations of the step kernels.  Built with reciprocal
// division (x / y as x * rcp(y) with Newton steps, ~1 ulp) and the hardware FP32 divide and
// sqrt (~2.5 ulp): the correctly rounded divide sequences were a sixth to a third of the
// kernels' instructions, and the difference is far below the FP32 rounding these builds
// already carry.
#include "ibis_step_kernels.hpp"

namespace ibis {

int launch_mixed(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
                 int64_t slot_pitch, const double* table, const IbisParams& P,
                 const StepConst& K, int64_t ncell, int copy_in, hipStream_t st) {
  return launch<float, float, double>(in, out, ws, slot, slot_pitch, table, P, K, ncell, copy_in, st);
}

int launch_f32(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
               int64_t slot_pitch, const double* table, const IbisParams& P,
               const StepConst& K, int64_t ncell, int copy_in, hipStream_t st) {
  return launch<float, float, float>(in, out, ws, slot, slot_pitch, table, P, K, ncell, copy_in, st);
}

int launch_ext_mixed(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
                     int64_t slot_pitch, const double* tables, const IbisParams& P,
                     const StepConst& K, const StepExt& E, int64_t ncell, int copy_in,
                     hipStream_t st) {
  return launch_ext<float, float, double>(in, out, ws, slot, slot_pitch, tables, P, K, E, ncell,
                                          copy_in, st);
}

int launch_ext_f32(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
                   int64_t slot_pitch, const double* tables, const IbisParams& P,
                   const StepConst& K, const StepExt& E, int64_t ncell, int copy_in,
                   hipStream_t st) {
  return launch_ext<float, float, float>(in, out, ws, slot, slot_pitch, tables, P, K, E, ncell,
                                         copy_in, st);
}

}  // namespace ibis
