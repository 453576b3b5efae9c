/*
 * C ABI of the daily step: the once-a-day work of the production driver around its hourly
 * steps (python-port/ibis/timestepping/loops_gswp3.py:382-583), on device, so that a month
 * (days of 24 hourly steps between a day begin and a day end) runs without host round trips.
 *
 *   ibis_day_begin  leaf phenology (python-port/ibis/physics/phenology.py:94-287) from the
 *                   10-day means and the day's mean temperature: cold and drought triggers,
 *                   growing degree-days, plant-type fractions, canopy LAI and lower-canopy
 *                   height written into the state; the daily sums of the accumulators
 *                   (adrain adsnow adaet a10ancub a10ancls a10ancl3 a10ancl4) reset
 *                   (loops_gswp3.py:403-416).
 *   ibis_day_end    10-day means from the day's sums (weight exp(-1/10) per day,
 *                   loops_gswp3.py:470-477), annual precipitation and AET, the decomposition
 *                   climate factors (:494-553), soil biogeochemistry nsoilcall times
 *                   (physics/biogeochem.py:23-278; tco2mic written to the vegetation extras,
 *                   where the hourly accumulators read it), annual heterotrophic and root
 *                   respiration (:577-585).
 *
 * Day arena: IBIS_NDAYPLANE planes of `pitch` cells in the state's element type (float for
 * IBIS_PREC_MIXED / F32, double for F64), fields below.  The day's mean temperature is plane
 * 0 (td) of the daily-weather arena of ibis_forcing_diurnal.  Default configuration of the
 * reference: scalar Lloyd-Taylor E0, no SOM-specific sensitivity, positional plant-type
 * habits; the seasonal states and photoperiod senescence are optional (IbisDayParams).
 * Stream-ordered; allocates nothing.
 */
#ifndef IBIS_DAILY_H
#define IBIS_DAILY_H

#include <stdint.h>
#include "ibis_driver.h"
#include "ibis_step.h"

#ifdef __cplusplus
extern "C" {
#endif

enum IbisDayPlane {
  IBIS_DAY_TEMPU = 0, IBIS_DAY_TEMPL, IBIS_DAY_AGDDU, IBIS_DAY_AGDDL,
  IBIS_DAY_DROPU, IBIS_DAY_DROPLS, IBIS_DAY_DROPL3, IBIS_DAY_DROPL4,
  IBIS_DAY_PLAI,                          /* 12: potential LAI per plant type */
  IBIS_DAY_TC = IBIS_DAY_PLAI + IBIS_NPFT,  /* coldest-month temperature (C) */
  IBIS_DAY_TCO2ROOT, IBIS_DAY_TNMIN,
  IBIS_DAY_CARBON,                        /* 24: CarbonState fields in dataclass order */
  IBIS_DAY_A10TD = IBIS_DAY_CARBON + 24,  /* 10-day means */
  IBIS_DAY_A10ANCUB, IBIS_DAY_A10ANCLS, IBIS_DAY_A10ANCL3, IBIS_DAY_A10ANCL4,
  IBIS_DAY_AYPRCP, IBIS_DAY_AYAET, IBIS_DAY_AYCO2MIC, IBIS_DAY_AYCO2ROOT,
  IBIS_DAY_STOREDN, IBIS_DAY_YRLEACH, IBIS_DAY_YNLEACH,
  IBIS_NDAYPLANE
};
/* offsets inside IBIS_DAY_CARBON (python-port/ibis/state/carbon.py) */
enum IbisCarbon {
  IBIS_C_CLITLM = 0, IBIS_C_CLITLS, IBIS_C_CLITLL, IBIS_C_CLITRM, IBIS_C_CLITRS,
  IBIS_C_CLITRL, IBIS_C_CLITWM, IBIS_C_CLITWS, IBIS_C_CLITWL, IBIS_C_CSOIPAS,
  IBIS_C_CSOISLON, IBIS_C_CSOISLOP, IBIS_C_DECOMPL, IBIS_C_DECOMPS, IBIS_C_FALLL,
  IBIS_C_FALLR, IBIS_C_FALLW, IBIS_C_CDISTURB, IBIS_C_TOTCMIC, IBIS_C_TOTLIT,
  IBIS_C_TOTALIT, IBIS_C_TOTRLIT, IBIS_C_TOTCSOI, IBIS_C_TOTFALL, IBIS_NCARBON
};

typedef struct IbisDayParams {
  /* PhenoParams (physics/phenology.py:20-85) */
  double tthresh_u_offset, gthresh_u_c, gdd_thresh_u, gdd_ramp_u;
  double tthresh_l_c, gthresh_l_c, gdd_thresh_l, gdd_ramp_l;
  double drop_thresh, drop_floor, ddays, drop_stress_thresh, tthresh_u_abs;
  double epsilon, tmelt;
  /* BGCParams (io/parameters.py:176-250) */
  double lig_frac, fbsom, effac, fmax, rconst, cnleaf, cnroot, cnwood, cnr[8];
  double klm, kls, kll, krm, krs, krl, kwm, kws, kwl, kbn, kbp, knb, kns, kpb, kps, ksb;
  double ylm, yls, yll, yrm, yrs, yrl, ywm, yws, ywl, ybn, ybp, yps, yns, ysb, ypb, ynb;
  double lt_e0, wfps_opt, k_scale;
  double hsoi[IBIS_NSOILAY + 1];
  double ndaypy, day_sec;
  int32_t nsoilcall, pad0;
  /* Seasonal states (python-port/ibis/state/seasonal.py:36-58; loops_gswp3.py:418-440).
   * With `seasonal` set, day begin first advances them from the day's mean temperature T:
   * the acclimation state S += (T - S) / acclim_tau, and per canopy the senescence sum, which
   * adds max(sen_tb - T, 0) while day length (from lat_rad and jday) is decreasing and below
   * sen_pstart and resets otherwise; with sen_photoperiod_u / _l a sum >= sen_ycrit forces the
   * canopy's cold branch and degree-day reset (physics/phenology.py:148-150, :170-172).
   * seasonal: 3 planes of the day arena's pitch (acclim_x deg C, ssen_u, ssen_l deg C day) in
   * the state's element type, device memory; lat_rad: ncell doubles, device memory. */
  double acclim_tau, sen_pstart_u, sen_tb_u, sen_ycrit_u, sen_pstart_l, sen_tb_l, sen_ycrit_l;
  int32_t sen_photoperiod_u, sen_photoperiod_l;
  int32_t jday, pad1;        /* day of year of this day, 1-based */
  void* seasonal;
  const double* lat_rad;
  /* Soil-water stress accumulator of the tree-establishment gate (state/stress_accum.py:15-26;
   * loops_gswp3.py:587-599).  With `stress` set, day end adds the canopies' stress factors
   * soil.stresstu / stresstl (after the day's last hour) to the sums and 1 to the day count on
   * growing-season days, td > tmelt + stress_accum_tmin (deg C; -100: every day).  stress: 3
   * planes of the day arena's pitch (aystressu, aystressl, nday) in the state's element type,
   * device memory. */
  double stress_accum_tmin;
  void* stress;
} IbisDayParams;

/* Day begin.  state: hourly-step arena (vegetation frac / lai / ztop / zbot written, fu fl
 * soil.stresstu stresstl read); day: day arena; daily: daily weather (plane 0 = td, K);
 * acc: the accumulators of ibis_accumulate (daily sums and their compensations zeroed). */
int ibis_day_begin(const IbisArena* state, const IbisArena* day, const IbisArena* daily,
                   const IbisArena* acc, const IbisDayParams* p, int64_t ncell, int flags,
                   void* hip_stream);

/* Day end.  Reads the day's sums from acc and soil tsoi / wsoi / wisoi / sand / clay from the
 * state; updates the day arena and writes tco2mic into vegx (IBIS_VEGX_TCO2MIC). */
int ibis_day_end(const IbisArena* state, const IbisArena* day, const IbisArena* daily,
                 const IbisArena* acc, const IbisArena* vegx, const IbisDayParams* p,
                 int64_t ncell, int flags, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* IBIS_DAILY_H */
