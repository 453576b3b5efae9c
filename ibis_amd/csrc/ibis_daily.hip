// HIP kernels and C ABI of the daily step (include/ibis_daily.h): one thread per cell,
// 256-thread blocks, every plane read and written with unit stride across the wavefront.
#include <hip/hip_runtime.h>

#include <stdio.h>

#include "ibis_daily_cell.hpp"
#include "ibis_error.hpp"

namespace ibis {

constexpr int kDayBlock = 256;

template <class S>
__global__ __launch_bounds__(kDayBlock) void day_begin_kernel(DayView<S> v, int64_t ncell,
                                                              IbisDayParams p) {
  const int64_t i = (int64_t)blockIdx.x * kDayBlock + threadIdx.x;
  if (i >= ncell) return;
  v.i = i;
  day_begin_cell(v, p);
}

template <class S>
__global__ __launch_bounds__(kDayBlock) void day_end_kernel(DayView<S> v, S* vegx,
                                                            int64_t vpitch, int64_t ncell,
                                                            IbisDayParams p) {
  const int64_t i = (int64_t)blockIdx.x * kDayBlock + threadIdx.x;
  if (i >= ncell) return;
  v.i = i;
  day_end_cell(v, vegx, vpitch, p);
}

static int day_check(const IbisArena* state, const IbisArena* day, const IbisArena* daily,
                     const IbisArena* acc, const IbisDayParams* p, int64_t ncell, int flags,
                     const char* what) {
  err_buf()[0] = 0;
  char msg[256];
  auto bad = [&](const char* why) {
    snprintf(msg, sizeof msg, "%s: %s", what, why);
    return fail(IBIS_E_ARG, msg);
  };
  if (!state || !state->planes || !day || !day->planes || !daily || !daily->planes || !acc ||
      !acc->planes || !p)
    return bad("null pointer argument");
  if (ncell < 0) return bad("negative ncell");
  if (state->pitch < ncell || day->pitch < ncell || daily->pitch < ncell || acc->pitch < ncell)
    return bad("every arena's pitch must be >= ncell");
  const int prec = flags & IBIS_PREC_MASK;
  if ((flags & ~IBIS_PREC_MASK) != 0 || prec > 2) {
    snprintf(msg, sizeof msg, "%s: unknown flag bits or precision", what);
    return fail(IBIS_E_FLAGS, msg);
  }
  if (!(p->ddays > 0.0) || !(p->ndaypy > 0.0) || !(p->day_sec > 0.0) || p->nsoilcall < 0)
    return bad("ddays, ndaypy and day_sec must be positive and nsoilcall >= 0");
  return IBIS_OK;
}

template <class S>
static DayView<S> day_view(const IbisArena* state, const IbisArena* day, const IbisArena* daily,
                           const IbisArena* acc) {
  DayView<S> v;
  v.st = (S*)state->planes;
  v.day = (S*)day->planes;
  v.daily = (const S*)daily->planes;
  v.acc = (S*)acc->planes;
  v.sp = state->pitch;
  v.dp = day->pitch;
  v.wp = daily->pitch;
  v.ap = acc->pitch;
  v.i = 0;
  return v;
}

static int launched(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(err_buf(), 512, "%s: launch failed: %s", what, hipGetErrorString(e));
    return IBIS_E_LAUNCH;
  }
  return IBIS_OK;
}

}  // namespace ibis

extern "C" {

int ibis_day_begin(const IbisArena* state, const IbisArena* day, const IbisArena* daily,
                   const IbisArena* acc, const IbisDayParams* p, int64_t ncell, int flags,
                   void* hip_stream) {
  using namespace ibis;
  const int rc = day_check(state, day, daily, acc, p, ncell, flags, "ibis_day_begin");
  if (rc != IBIS_OK) return rc;
  if (p->seasonal && (!p->lat_rad || !(p->acclim_tau != 0.0)))
    return fail(IBIS_E_ARG, "ibis_day_begin: the seasonal states need lat_rad and a nonzero "
                            "acclim_tau");
  if (ncell == 0) return IBIS_OK;
  const dim3 grid((unsigned)((ncell + kDayBlock - 1) / kDayBlock)), block(kDayBlock);
  hipStream_t s = (hipStream_t)hip_stream;
  if ((flags & IBIS_PREC_MASK) == IBIS_PREC_F64)
    hipLaunchKernelGGL(day_begin_kernel<double>, grid, block, 0, s,
                       day_view<double>(state, day, daily, acc), ncell, *p);
  else
    hipLaunchKernelGGL(day_begin_kernel<float>, grid, block, 0, s,
                       day_view<float>(state, day, daily, acc), ncell, *p);
  return launched("ibis_day_begin");
}

int ibis_day_end(const IbisArena* state, const IbisArena* day, const IbisArena* daily,
                 const IbisArena* acc, const IbisArena* vegx, const IbisDayParams* p,
                 int64_t ncell, int flags, void* hip_stream) {
  using namespace ibis;
  int rc = day_check(state, day, daily, acc, p, ncell, flags, "ibis_day_end");
  if (rc != IBIS_OK) return rc;
  if (!vegx || !vegx->planes || vegx->pitch < ncell)
    return fail(IBIS_E_ARG, "ibis_day_end: null vegetation extras or pitch < ncell");
  if (ncell == 0) return IBIS_OK;
  const dim3 grid((unsigned)((ncell + kDayBlock - 1) / kDayBlock)), block(kDayBlock);
  hipStream_t s = (hipStream_t)hip_stream;
  if ((flags & IBIS_PREC_MASK) == IBIS_PREC_F64)
    hipLaunchKernelGGL(day_end_kernel<double>, grid, block, 0, s,
                       day_view<double>(state, day, daily, acc), (double*)vegx->planes,
                       vegx->pitch, ncell, *p);
  else
    hipLaunchKernelGGL(day_end_kernel<float>, grid, block, 0, s,
                       day_view<float>(state, day, daily, acc), (float*)vegx->planes,
                       vegx->pitch, ncell, *p);
  return launched("ibis_day_end");
}

}  // extern "C"
