// HIP kernels and C ABI of the hourly step (include/ibis_step.h).
//
// One thread per land cell, 256 threads (four wave64) per block.  The plant-type table
// (52 rows) is staged once per block into LDS and each thread gathers its 12 slot values
// from it; every state plane is read and written with unit stride across the wavefront.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "ibis_step_cell.hpp"

namespace ibis {

constexpr int kBlock = 256;

template <class S, class R, class D>
__global__ __launch_bounds__(kBlock) void step_kernel(const S* in, S* out, int64_t pitch, int64_t ncell,
                                                      const int8_t* slot, int64_t slot_pitch,
                                                      const double* table, IbisParams P,
                                                      StepConst K, int copy_in) {
  __shared__ R trow[IBIS_TYPE_ROWS];
  for (int t = threadIdx.x; t < IBIS_TYPE_ROWS; t += blockDim.x) trow[t] = R(table[t]);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= ncell) return;
  R vrow[IBIS_NPFT];
#pragma unroll
  for (int j = 0; j < IBIS_NPFT; ++j) {
    const int s = slot ? (int)slot[(int64_t)j * slot_pitch + i] : -1;
    const int row = (s >= 0 && s < IBIS_NTYPE) ? s : IBIS_NTYPE + j;
    vrow[j] = trow[row];
  }
  if (copy_in) {
#define X(F, W, OFF, ROLE)                                                      \
  if ((ROLE) == IBIS_ROLE_IN) {                                                 \
    for (int j = 0; j < (W); ++j)                                               \
      out[(int64_t)((OFF) + j) * pitch + i] = in[(int64_t)((OFF) + j) * pitch + i]; \
  }
    IBIS_FIELD_TABLE
#undef X
  }
  Cell<S> c{in, out, pitch, i};
  step_cell<S, R, D>(c, vrow, P, K);
}

static thread_local char g_err[512] = "";

static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

template <class S, class R, class D>
static int launch(const IbisArena* in, const IbisArena* out, const int8_t* slot,
                  int64_t slot_pitch, const double* table, const IbisParams& P,
                  const StepConst& K, int64_t ncell, int copy_in, hipStream_t st) {
  const int64_t nblk = (ncell + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((step_kernel<S, R, D>), dim3((unsigned)nblk), dim3(kBlock), 0, st,
                     (const S*)in->planes, (S*)out->planes, in->pitch,
                     ncell, slot, slot_pitch, table, P, K, copy_in);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "ibis_step: launch failed: %s", hipGetErrorString(e));
    return IBIS_E_LAUNCH;
  }
  return IBIS_OK;
}

}  // namespace ibis

extern "C" {

const char* ibis_last_error(void) { return ibis::g_err; }

const char* ibis_version(void) { return "ibis_amd 0.1 (gfx950)"; }

int ibis_step(const IbisArena* in, const IbisArena* out, const int8_t* slot_type,
              int64_t slot_pitch, const double* type_table, const IbisParams* params,
              double dtime, double nytimes, int64_t ncell, int flags, void* hip_stream) {
  using namespace ibis;
  g_err[0] = 0;
  if (!in || !out || !params || !type_table || !in->planes || !out->planes)
    return fail(IBIS_E_ARG, "ibis_step: null pointer argument");
  if (ncell < 0) return fail(IBIS_E_ARG, "ibis_step: negative ncell");
  if (in->pitch < ncell || out->pitch != in->pitch)
    return fail(IBIS_E_ARG, "ibis_step: pitch must be >= ncell and equal for in and out");
  if (slot_type && slot_pitch < ncell)
    return fail(IBIS_E_ARG, "ibis_step: slot_pitch must be >= ncell");
  if (!(dtime > 0.0)) return fail(IBIS_E_ARG, "ibis_step: dtime must be positive");
  const int prec = flags & IBIS_PREC_MASK;
  if ((flags & ~(IBIS_PREC_MASK | IBIS_FLAG_BEDROCK | IBIS_FLAG_COPY_IN)) != 0 || prec > 2)
    return fail(IBIS_E_FLAGS, "ibis_step: unknown flag bits or precision");
  const IbisParams& P = *params;
  if (flags & IBIS_FLAG_BEDROCK) {
    for (int k = 0; k < IBIS_NBEDROCK; ++k)
      if (!(P.bed_hsoi[k] > 0.0))
        return fail(IBIS_E_BEDROCK, "ibis_step: bedrock layer thicknesses must be positive");
  }
  if (ncell == 0) return IBIS_OK;
  StepConst K;
  K.dtime = dtime;
  K.tfac = 1.0 / pow(0.999, P.cappa);
  K.zweight = exp(-1.0 / (10.0 * 86400.0 / dtime));
  K.nytimes_new = nytimes + 1.0;
  K.bedrock = (flags & IBIS_FLAG_BEDROCK) ? 1 : 0;
  const int copy_in = (flags & IBIS_FLAG_COPY_IN) ? 1 : 0;
  hipStream_t st = (hipStream_t)hip_stream;
  switch (prec) {
    case IBIS_PREC_MIXED:
      return launch<float, float, double>(in, out, slot_type, slot_pitch, type_table, P, K, ncell,
                                          copy_in, st);
    case IBIS_PREC_F32:
      return launch<float, float, float>(in, out, slot_type, slot_pitch, type_table, P, K, ncell,
                                         copy_in, st);
    default:
      return launch<double, double, double>(in, out, slot_type, slot_pitch, type_table, P, K,
                                            ncell, copy_in, st);
  }
}

}  // extern "C"
