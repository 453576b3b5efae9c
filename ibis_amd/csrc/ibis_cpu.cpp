// The hourly step on host threads (include/ibis_cpu.h): the three phases of ibis_step over
// contiguous chunks of cells, one chunk per thread, with the kernels' per-cell source.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "ibis_error.hpp"
#include "ibis_step_cell.hpp"

namespace {

template <class S, class R, class D>
void run_chunk(const S* in, S* out, int64_t pitch, int64_t c0, int64_t c1, const int8_t* slot,
               int64_t slot_pitch, const double* table, const IbisParams& P,
               const ibis::StepConst& K, int copy_in, D* wd, R* wr) {
  auto work = [&](int64_t i) { return ibis::Work<R, D>{wd, wr, pitch, i}; };
  for (int64_t i = c0; i < c1; ++i) {
    if (copy_in) {
#define X(F, W, OFF, ROLE)                                                            \
  if ((ROLE) == IBIS_ROLE_IN)                                                         \
    for (int j = 0; j < (W); ++j)                                                     \
      out[(int64_t)((OFF) + j) * pitch + i] = in[(int64_t)((OFF) + j) * pitch + i];
      IBIS_FIELD_TABLE
#undef X
    }
    const ibis::Cell<S> c{in, out, pitch, i};
    ibis::surface_cell<S, R, D>(c, work(i), P, K);
  }
  for (int64_t i = c0; i < c1; ++i) {
    R vrow[IBIS_NPFT];
    for (int j = 0; j < IBIS_NPFT; ++j) {
      const int s = slot ? (int)slot[(int64_t)j * slot_pitch + i] : -1;
      vrow[j] = R(table[(s >= 0 && s < IBIS_NTYPE) ? s : IBIS_NTYPE + j]);
    }
    const ibis::Cell<S> c{in, out, pitch, i};
    ibis::canopy_cell<S, R, D>(c, work(i), vrow, P, K);
  }
  for (int64_t i = c0; i < c1; ++i) {
    const ibis::Cell<S> c{in, out, pitch, i};
    ibis::ground_cell<S, R, D>(c, work(i), P, K);
  }
}

template <class S, class R, class D>
void run(const IbisArena* in, const IbisArena* out, const int8_t* slot, int64_t slot_pitch,
         const double* table, const IbisParams& P, const ibis::StepConst& K, int64_t ncell,
         int copy_in, int nthreads) {
  const int64_t pitch = in->pitch;
  std::vector<D> wd((size_t)ibis::WD_N * (size_t)pitch);
  std::vector<R> wr((size_t)ibis::WR_N * (size_t)pitch);
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(nthreads, ncell));
  const int64_t chunk = (ncell + nt - 1) / nt;
  std::vector<std::thread> pool;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t c0 = t * chunk, c1 = std::min(ncell, c0 + chunk);
    if (c0 >= c1) break;
    pool.emplace_back(run_chunk<S, R, D>, (const S*)in->planes, (S*)out->planes, pitch, c0, c1,
                      slot, slot_pitch, table, std::cref(P), std::cref(K), copy_in, wd.data(),
                      wr.data());
  }
  for (auto& th : pool) th.join();
}

}  // namespace

extern "C" int ibis_step_cpu(const IbisArena* in, const IbisArena* out, const int8_t* slot_type,
                             int64_t slot_pitch, const double* type_table,
                             const IbisParams* params, double dtime, double nytimes,
                             int64_t ncell, int flags, int nthreads) {
  using namespace ibis;
  err_buf()[0] = 0;
  if (!in || !out || !params || !type_table || !in->planes || !out->planes)
    return fail(IBIS_E_ARG, "ibis_step_cpu: null pointer argument");
  if (ncell < 0 || in->pitch < ncell || out->pitch != in->pitch)
    return fail(IBIS_E_ARG, "ibis_step_cpu: pitch must be >= ncell and equal for in and out");
  if (slot_type && slot_pitch < ncell)
    return fail(IBIS_E_ARG, "ibis_step_cpu: slot_pitch must be >= ncell");
  if (!(dtime > 0.0)) return fail(IBIS_E_ARG, "ibis_step_cpu: dtime must be positive");
  const int prec = flags & IBIS_PREC_MASK;
  if ((flags & ~(IBIS_PREC_MASK | IBIS_FLAG_BEDROCK | IBIS_FLAG_COPY_IN)) != 0 || prec > 2)
    return fail(IBIS_E_FLAGS, "ibis_step_cpu: unknown flag bits or precision");
  const IbisParams& P = *params;
  if (flags & IBIS_FLAG_BEDROCK) {
    for (int k = 0; k < IBIS_NBEDROCK; ++k)
      if (!(P.bed_hsoi[k] > 0.0))
        return fail(IBIS_E_BEDROCK, "ibis_step_cpu: bedrock layer thicknesses must be positive");
  }
  if (ncell == 0) return IBIS_OK;
  StepConst K;
  K.dtime = dtime;
  K.tfac = 1.0 / pow(0.999, P.cappa);
  K.zweight = exp(-1.0 / (10.0 * 86400.0 / dtime));
  K.nytimes_new = nytimes + 1.0;
  K.bedrock = (flags & IBIS_FLAG_BEDROCK) ? 1 : 0;
  const int copy_in = (flags & IBIS_FLAG_COPY_IN) ? 1 : 0;
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  switch (prec) {
    case IBIS_PREC_MIXED:
      run<float, float, double>(in, out, slot_type, slot_pitch, type_table, P, K, ncell, copy_in,
                                nthreads);
      break;
    case IBIS_PREC_F32:
      run<float, float, float>(in, out, slot_type, slot_pitch, type_table, P, K, ncell, copy_in,
                               nthreads);
      break;
    default:
      run<double, double, double>(in, out, slot_type, slot_pitch, type_table, P, K, ncell,
                                  copy_in, nthreads);
  }
  return IBIS_OK;
}
