// HIP kernels and C ABI of the hourly step (include/ibis_step.h).
//
// One thread per land cell.  The step runs as three kernels in stream order -- surface,
// canopy, ground (ibis_step_cell.hpp) -- so that each is register-allocated on its own; the
// values they hand on live in a plane-major workspace.  Every state and workspace plane is
// read and written with unit stride across the wavefront.  The canopy kernel stages the
// plant-type table (52 rows) in LDS once per block and gathers each cell's 12 slot values
// from it.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "ibis_error.hpp"
#include "ibis_step_cell.hpp"

namespace ibis {

constexpr int kBlock = 256;
// Waves per SIMD each kernel is compiled for.  95,333 cells are 1,490 waves on 1,024 SIMDs,
// so a kernel held to one wave per SIMD runs its waves in two rounds; surface and ground fit
// two waves at the cost of a few spilled registers, which measures faster.  The canopy
// kernel's three passes spill far more at two waves and stay at one.
constexpr int kWavesSurface = 2, kWavesCanopy = 1, kWavesGround = 2;

template <class R, class D>
__device__ IBIS_INLINE Work<R, D> work_at(void* ws, int64_t pitch, int64_t i) {
  D* d = (D*)ws;
  R* r = (R*)(d + (int64_t)WD_N * pitch);
  return Work<R, D>{d, r, pitch, i};
}

template <class S, class R, class D>
__global__ __launch_bounds__(kBlock, kWavesSurface) void surface_kernel(
    const S* in, S* out, void* ws, int64_t pitch, int64_t ncell, IbisParams P, StepConst K,
    int copy_in) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= ncell) return;
  if (copy_in) {
#define X(F, W, OFF, ROLE)                                                      \
  if ((ROLE) == IBIS_ROLE_IN) {                                                 \
    for (int j = 0; j < (W); ++j)                                               \
      out[(int64_t)((OFF) + j) * pitch + i] = in[(int64_t)((OFF) + j) * pitch + i]; \
  }
    IBIS_FIELD_TABLE
#undef X
  }
  const Cell<S> c{in, out, pitch, i};
  surface_cell<S, R, D>(c, work_at<R, D>(ws, pitch, i), P, K);
}

template <class S, class R, class D>
__global__ __launch_bounds__(kBlock, kWavesCanopy) void canopy_kernel(
    const S* in, S* out, void* ws, int64_t pitch, int64_t ncell, const int8_t* slot,
    int64_t slot_pitch, const double* table, IbisParams P, StepConst K) {
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
  const Cell<S> c{in, out, pitch, i};
  canopy_cell<S, R, D>(c, work_at<R, D>(ws, pitch, i), vrow, P, K);
}

template <class S, class R, class D>
__global__ __launch_bounds__(kBlock, kWavesGround) void ground_kernel(
    const S* in, S* out, void* ws, int64_t pitch, int64_t ncell, IbisParams P, StepConst K) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= ncell) return;
  const Cell<S> c{in, out, pitch, i};
  ground_cell<S, R, D>(c, work_at<R, D>(ws, pitch, i), P, K);
}

template <class S, class R, class D>
static int launch(const IbisArena* in, const IbisArena* out, void* ws, const int8_t* slot,
                  int64_t slot_pitch, const double* table, const IbisParams& P,
                  const StepConst& K, int64_t ncell, int copy_in, hipStream_t st) {
  const dim3 grid((unsigned)((ncell + kBlock - 1) / kBlock)), block(kBlock);
  const S* pin = (const S*)in->planes;
  S* pout = (S*)out->planes;
  const int64_t pitch = in->pitch;
  hipLaunchKernelGGL((surface_kernel<S, R, D>), grid, block, 0, st, pin, pout, ws, pitch, ncell,
                     P, K, copy_in);
  hipLaunchKernelGGL((canopy_kernel<S, R, D>), grid, block, 0, st, pin, pout, ws, pitch, ncell,
                     slot, slot_pitch, table, P, K);
  hipLaunchKernelGGL((ground_kernel<S, R, D>), grid, block, 0, st, pin, pout, ws, pitch, ncell,
                     P, K);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(err_buf(), 512, "ibis_step: launch failed: %s", hipGetErrorString(e));
    return IBIS_E_LAUNCH;
  }
  return IBIS_OK;
}

static int64_t workspace_bytes(int64_t pitch, int prec) {
  switch (prec) {
    case IBIS_PREC_MIXED: return work_bytes(pitch, (int)sizeof(float), (int)sizeof(double));
    case IBIS_PREC_F32: return work_bytes(pitch, (int)sizeof(float), (int)sizeof(float));
    default: return work_bytes(pitch, (int)sizeof(double), (int)sizeof(double));
  }
}

}  // namespace ibis

extern "C" {

const char* ibis_last_error(void) { return ibis::err_buf(); }

const char* ibis_version(void) { return "ibis_amd 0.2 (gfx950, three-phase step)"; }

int64_t ibis_step_workspace_bytes(int64_t pitch, int flags) {
  const int prec = flags & IBIS_PREC_MASK;
  if (pitch < 0 || prec > 2) return IBIS_E_ARG;
  return ibis::workspace_bytes(pitch, prec);
}

int ibis_step(const IbisArena* in, const IbisArena* out, void* workspace,
              int64_t workspace_bytes, const int8_t* slot_type, int64_t slot_pitch,
              const double* type_table, const IbisParams* params, double dtime, double nytimes,
              int64_t ncell, int flags, void* hip_stream) {
  using namespace ibis;
  err_buf()[0] = 0;
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
  if (!workspace || workspace_bytes < ibis::workspace_bytes(in->pitch, prec))
    return fail(IBIS_E_ARG, "ibis_step: workspace missing or smaller than "
                            "ibis_step_workspace_bytes(pitch, flags)");
  if (((uintptr_t)workspace & 255u) != 0)
    return fail(IBIS_E_ARG, "ibis_step: workspace must be 256-byte aligned");
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
      return launch<float, float, double>(in, out, workspace, slot_type, slot_pitch, type_table,
                                          P, K, ncell, copy_in, st);
    case IBIS_PREC_F32:
      return launch<float, float, float>(in, out, workspace, slot_type, slot_pitch, type_table,
                                         P, K, ncell, copy_in, st);
    default:
      return launch<double, double, double>(in, out, workspace, slot_type, slot_pitch,
                                            type_table, P, K, ncell, copy_in, st);
  }
}

}  // extern "C"
