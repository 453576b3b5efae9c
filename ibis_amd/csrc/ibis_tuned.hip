// C ABI of the parameter-estimation step (include/ibis_adjoint.h): argument checks, launch
// constants and dispatch to the per-precision kernel sets (ibis_tuned_fast.hip,
// ibis_tuned_f64.hip).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>

#include "ibis_tuned_kernels.hpp"

namespace ibis {

int tuned_launch_mixed(void* a, void* ws, const TunedArgs& g, const IbisParams& base,
                       const StepConst& K, const MisfitArgs& m, const float* target, float* flux,
                       double* loss, bool tangent, hipStream_t st);
int tuned_launch_f64(void* a, void* ws, const TunedArgs& g, const IbisParams& base,
                     const StepConst& K, const MisfitArgs& m, const float* target, float* flux,
                     double* loss, bool tangent, hipStream_t st);

static int64_t tuned_work_bytes(int64_t pitch, int prec, bool tangent) {
  const int r = prec == IBIS_PREC_F64 ? 8 : 4;
  const int k = tangent ? 2 : 1;
  return work_bytes(pitch, r * k, 8 * k);
}

}  // namespace ibis

extern "C" {

int64_t ibis_tuned_workspace_bytes(int64_t pitch, int flags) {
  const int prec = flags & IBIS_PREC_MASK;
  if (pitch < 0 || (prec != IBIS_PREC_MIXED && prec != IBIS_PREC_F64) ||
      (flags & ~(IBIS_PREC_MASK | IBIS_FLAG_TANGENT | IBIS_FLAG_BEDROCK)) != 0)
    return IBIS_E_ARG;
  return ibis::tuned_work_bytes(pitch, prec, (flags & IBIS_FLAG_TANGENT) != 0);
}

int ibis_tuned_hour(const IbisArena* state, void* workspace, int64_t workspace_bytes,
                    const int8_t* slot_type, const double* lat_deg, const double* tune,
                    const int32_t* row_pft, const IbisParams* base, const IbisTuned* run,
                    const float* target, float* flux, double* loss, int flags,
                    void* hip_stream) {
  using namespace ibis;
  err_buf()[0] = 0;
  if (!state || !state->planes || !workspace || !slot_type || !lat_deg || !tune || !row_pft ||
      !base || !run)
    return fail(IBIS_E_ARG, "ibis_tuned_hour: null pointer argument");
  const int prec = flags & IBIS_PREC_MASK;
  const bool tangent = (flags & IBIS_FLAG_TANGENT) != 0;
  if ((flags & ~(IBIS_PREC_MASK | IBIS_FLAG_TANGENT | IBIS_FLAG_BEDROCK)) != 0 ||
      (prec != IBIS_PREC_MIXED && prec != IBIS_PREC_F64))
    return fail(IBIS_E_FLAGS, "ibis_tuned_hour: unknown flag bits, or precision other than "
                              "mixed or FP64");
  const IbisTuned& r = *run;
  if (r.nsite <= 0 || r.npair <= 0 || r.ndir <= 0 || r.npair % r.nsite != 0 || r.hour < 0 ||
      r.site0 < 0)
    return fail(IBIS_E_ARG, "ibis_tuned_hour: nsite, npair, ndir must be positive, npair a "
                            "multiple of nsite, hour and site0 >= 0");
  if (tangent ? (r.dir0 < 0 || r.dir0 + r.ndir > IBIS_NTUNE) : (r.dir0 != -1 || r.ndir != 1))
    return fail(IBIS_E_ARG, "ibis_tuned_hour: tangent mode needs 0 <= dir0 and dir0 + ndir <= "
                            "IBIS_NTUNE; value mode needs dir0 = -1 and ndir = 1");
  const int64_t nthread = r.ndir * r.npair;
  if (state->pitch < nthread)
    return fail(IBIS_E_ARG, "ibis_tuned_hour: state pitch smaller than ndir * npair");
  if (!(r.dtime > 0.0)) return fail(IBIS_E_ARG, "ibis_tuned_hour: dtime must be positive");
  for (int k = 0; k < IBIS_NFLUX; ++k)
    if (!(r.flux_scale[k] > 0.0))
      return fail(IBIS_E_ARG, "ibis_tuned_hour: flux scales must be positive");
  if (workspace_bytes < tuned_work_bytes(state->pitch, prec, tangent))
    return fail(IBIS_E_ARG, "ibis_tuned_hour: workspace smaller than "
                            "ibis_tuned_workspace_bytes(pitch, flags)");
  if (((uintptr_t)workspace & 255u) != 0)
    return fail(IBIS_E_ARG, "ibis_tuned_hour: workspace must be 256-byte aligned");
  if (flags & IBIS_FLAG_BEDROCK) {
    for (int k = 0; k < IBIS_NBEDROCK; ++k)
      if (!(base->bed_hsoi[k] > 0.0))
        return fail(IBIS_E_BEDROCK, "ibis_tuned_hour: bedrock layer thicknesses must be positive");
  }
  StepConst K;
  K.dtime = r.dtime;
  K.tfac = 1.0 / pow(0.999, base->cappa);
  K.zweight = exp(-1.0 / (10.0 * 86400.0 / r.dtime));
  K.nytimes_new = r.nytimes + 1.0;
  K.bedrock = (flags & IBIS_FLAG_BEDROCK) ? 1 : 0;
  TunedArgs g;
  g.pitch = state->pitch;
  g.nsite = r.nsite;
  g.npair = r.npair;
  g.nthread = nthread;
  g.site0 = r.site0;
  g.dir0 = tangent ? r.dir0 : -1;
  g.jday0 = r.jday0;
  g.seed = r.seed;
  g.hour = r.hour;
  g.lat = lat_deg;
  g.tune = tune;
  g.slot = slot_type;
  g.row_pft = row_pft;
  MisfitArgs m;
  m.hvap = base->hvap;
  for (int k = 0; k < IBIS_NFLUX; ++k) m.inv_scale[k] = 1.0 / r.flux_scale[k];
  hipStream_t st = (hipStream_t)hip_stream;
  if (prec == IBIS_PREC_MIXED)
    return tuned_launch_mixed(state->planes, workspace, g, *base, K, m, target, flux, loss,
                              tangent, st);
  return tuned_launch_f64(state->planes, workspace, g, *base, K, m, target, flux, loss, tangent,
                          st);
}

}  // extern "C"
