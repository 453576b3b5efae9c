// Per-thread bodies of the parameter-estimation step (include/ibis_adjoint.h), shared by the
// HIP kernels (ibis_tuned_kernels.hpp) and the host build used by the CPU tests: forcing and
// the step's three phases with a pair's tunable parameter set, and the hourly flux misfit.
#pragma once

#include <type_traits>

#include "ibis_driver_cell.hpp"
#include "ibis_dual.hpp"
#include "ibis_step_cell.hpp"

namespace ibis {

// Per-launch arguments.
struct TunedArgs {
  int64_t pitch, nsite, npair, nthread, site0;
  int dir0, jday0;
  uint64_t seed;
  int64_t hour;
  const double* lat;
  const double* tune;
  const int8_t* slot;
  const int32_t* row_pft;
};

// Thread t steps pair t % npair (site pair % nsite) seeded along dir0 + t / npair.
struct TunedThread {
  int64_t t, pair, site;
  int dir;
};

IBIS_HD IBIS_INLINE TunedThread tuned_thread_of(const TunedArgs& g, int64_t t) {
  TunedThread th;
  th.t = t;
  const int64_t k = t / g.npair;
  th.pair = t - k * g.npair;
  th.site = th.pair % g.nsite;
  th.dir = g.dir0 + (int)k;
  return th;
}

template <class R, class D>
IBIS_HD IBIS_INLINE Work<R, D> tuned_work(void* ws, int64_t pitch, int64_t i) {
  D* d = (D*)ws;
  R* r = (R*)(d + (int64_t)WD_N * pitch);
  return Work<R, D>{d, r, pitch, i};
}

template <class S, class R, class D, class Q>
IBIS_HD IBIS_INLINE void tuned_surface(S* a, void* ws, const TunedArgs& g, const TunedThread& th,
                                       const IbisParams& base, const StepConst& K) {
  Forcing f;
  synthetic_forcing(f, g.lat[th.site], (double)a[(int64_t)IBIS_P_ATM_PSURF * g.pitch + th.t],
                    g.seed, (uint64_t)(g.site0 + th.site), g.jday0, g.hour);
  put_forcing(a, g.pitch, th.t, f);
  TunedParams<Q> P;
  tune_params(P, base, g.tune + th.pair, g.npair, th.dir);
  const Cell<S> c{a, a, g.pitch, th.t};
  surface_cell<S, R, D>(c, tuned_work<R, D>(ws, g.pitch, th.t), P, K);
}

template <class S, class R, class D, class Q>
IBIS_HD IBIS_INLINE void tuned_canopy(S* a, void* ws, const TunedArgs& g, const TunedThread& th,
                                      const IbisParams& base, const StepConst& K) {
  // Vmax of each slot: the pair's value for the plant-type position of the slot's row
  R vrow[IBIS_NPFT];
#pragma unroll
  for (int j = 0; j < IBIS_NPFT; ++j) {
    const int s = (int)g.slot[(int64_t)j * g.nsite + th.site];
    const int row = (s >= 0 && s < IBIS_NTYPE) ? s : IBIS_NTYPE + j;
    const int k = IBIS_TUNE_VMAX + g.row_pft[row];
    const double v = g.tune[(int64_t)k * g.npair + th.pair];
    if constexpr (is_dual<Q>::value) vrow[j] = R(Q(v, k == th.dir ? v : 0.0));
    else vrow[j] = R(v);
  }
  TunedParams<Q> P;
  tune_params(P, base, g.tune + th.pair, g.npair, th.dir);
  const Cell<S> c{a, a, g.pitch, th.t};
  canopy_cell<S, R, D>(c, tuned_work<R, D>(ws, g.pitch, th.t), vrow, P, K);
}

template <class S, class R, class D, class Q>
IBIS_HD IBIS_INLINE void tuned_ground(S* a, void* ws, const TunedArgs& g, const TunedThread& th,
                                      const IbisParams& base, const StepConst& K) {
  TunedParams<Q> P;
  tune_params(P, base, g.tune + th.pair, g.npair, th.dir);
  const Cell<S> c{a, a, g.pitch, th.t};
  ground_cell<S, R, D>(c, tuned_work<R, D>(ws, g.pitch, th.t), P, K);
}

// The hour's fluxes of a cell after the step: GPP (the driver's canopy gross assimilation,
// python-port/ibis/timestepping/loops_gswp3.py:239-252), LE and H.
template <class S, class L>
IBIS_HD IBIS_INLINE void tuned_fluxes(const S* a, int64_t pitch, int64_t i, double hvap,
                                      L f[IBIS_NFLUX]) {
  auto ld = [&](int p) { return L(a[(int64_t)p * pitch + i]); };
  L fr[IBIS_NPFT];
#pragma unroll
  for (int j = 0; j < IBIS_NPFT; ++j) fr[j] = ld(IBIS_P_VEGETATION_FRAC + j);
  const L fb = fr[0] + fr[1] + fr[2] + fr[4] + fr[6];
  const L fc = fr[3] + fr[5] + fr[7];
  const L smask = L(1) - ld(IBIS_P_SNOW_FI);
  const L gpp_u = (ld(IBIS_P_CANOPY_AG_UPPER) * fb + ld(IBIS_P_CANOPY_AG_UPPER + 1) * fc) *
                  ld(IBIS_P_VEGETATION_LAI + 1) * ld(IBIS_P_VEGETATION_FU);
  const L gpp_l = (ld(IBIS_P_CANOPY_AG_LOWER) * (fr[8] + fr[9]) +
                   ld(IBIS_P_CANOPY_AG_LOWER + 2) * fr[10] +
                   ld(IBIS_P_CANOPY_AG_LOWER + 1) * fr[11]) *
                  ld(IBIS_P_VEGETATION_LAI) * ld(IBIS_P_VEGETATION_FL) * smask;
  f[0] = gpp_u + gpp_l;
  f[1] = L(hvap) * (ld(IBIS_P_HYDROLOGY_GINVAP) + ld(IBIS_P_HYDROLOGY_GSUVAP) +
                    ld(IBIS_P_HYDROLOGY_GTRANS));
  f[2] = ld(IBIS_P_DIAG_FSENA);
}

struct MisfitArgs {
  double hvap;
  double inv_scale[IBIS_NFLUX];
};

// Store the fluxes of thread th and/or add its squared misfit (value and tangent; FP64).
template <class S>
IBIS_HD IBIS_INLINE void tuned_misfit(const S* a, const TunedArgs& g, const TunedThread& th,
                                      const MisfitArgs& m, const float* target, float* flux,
                                      double* loss) {
  using L = typename std::conditional<is_dual<S>::value, Dual<double>, double>::type;
  L f[IBIS_NFLUX];
  tuned_fluxes<S, L>(a, g.pitch, th.t, m.hvap, f);
  if (flux) {
#pragma unroll
    for (int k = 0; k < IBIS_NFLUX; ++k) {
      if constexpr (is_dual<L>::value) flux[(int64_t)k * g.pitch + th.t] = (float)f[k].v;
      else flux[(int64_t)k * g.pitch + th.t] = (float)f[k];
    }
  }
  if (target && loss) {
    L e2 = L(0);
#pragma unroll
    for (int k = 0; k < IBIS_NFLUX; ++k) {
      const L e = (f[k] - L((double)target[(int64_t)k * g.nsite + th.site])) * L(m.inv_scale[k]);
      e2 = e2 + e * e;
    }
    if constexpr (is_dual<L>::value) {
      loss[th.t] += e2.v;
      loss[g.pitch + th.t] += e2.d;
    } else {
      loss[th.t] += e2;
    }
  }
}

}  // namespace ibis
