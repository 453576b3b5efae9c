// Per-cell bodies of the daily step (include/ibis_daily.h), shared by the HIP kernels
// (ibis_daily.hip) and the host build.  Arithmetic in the state's element type S, in the
// reference's expression order; jnp.maximum / minimum / clip as the NaN-propagating mx / mn /
// clip of ibis_math.hpp.
#pragma once

#include "ibis_daily.h"
#include "ibis_driver_cell.hpp"
#include "ibis_math.hpp"

namespace ibis {

template <class S> struct DayView {
  S* st;          // hourly-step arena
  S* day;         // day arena
  const S* daily; // daily weather (plane 0: td)
  S* acc;         // accumulators: sums then compensations
  int64_t sp, dp, wp, ap, i;
  IBIS_HD IBIS_INLINE S& s(int p) const { return st[(int64_t)p * sp + i]; }
  IBIS_HD IBIS_INLINE S& d(int p) const { return day[(int64_t)p * dp + i]; }
  IBIS_HD IBIS_INLINE S& c(int k) const { return day[(int64_t)(IBIS_DAY_CARBON + k) * dp + i]; }
  IBIS_HD IBIS_INLINE S td() const { return daily[i]; }
  IBIS_HD IBIS_INLINE S sum(int k) const {        // an accumulator's value: sum - compensation
    return acc[(int64_t)k * ap + i] - acc[(int64_t)(IBIS_NACC + k) * ap + i];
  }
  IBIS_HD IBIS_INLINE void reset(int k) const {
    acc[(int64_t)k * ap + i] = S(0);
    acc[(int64_t)(IBIS_NACC + k) * ap + i] = S(0);
  }
};

// The day's daily sums, reset at day begin (loops_gswp3.py:403-416).
constexpr int kDailySums[7] = {IBIS_ACC_ADRAIN, IBIS_ACC_ADSNOW, IBIS_ACC_ADAET,
                               IBIS_ACC_A10ANCUB, IBIS_ACC_A10ANCLS, IBIS_ACC_A10ANCL3,
                               IBIS_ACC_A10ANCL4};

// ---- day begin: leaf phenology (physics/phenology.py:94-287) -------------------------------
template <class S>
IBIS_HD IBIS_INLINE S drop_update(S x, S anc, S stress, const IbisDayParams& p, S ddfac) {
  return (anc < S(p.drop_thresh) || stress < S(p.drop_stress_thresh))
             ? mx(S(p.drop_floor), x - ddfac)
             : mn(S(1), x + ddfac);
}

// Astronomical day length (h) at latitude lat (radians) on day of year doy
// (state/seasonal.py:53-58)
template <class S> IBIS_HD IBIS_INLINE S daylength_hours(S lat, S doy) {
  const S decl = S(0.40910518) * r_sin(S(2.0 * 3.14159265358979323846) * (doy - S(81)) / S(365));
  const S x = clip(-r_tan(lat) * r_tan(decl), S(-1), S(1));
  return S(24.0 / 3.14159265358979323846) * r_acos(x);
}

// Seasonal states, advanced once a day before phenology (loops_gswp3.py:418-440); returns
// the senescence triggers of the upper and lower canopy.
template <class S>
IBIS_HD IBIS_INLINE void seasonal_update(const DayView<S>& v, const IbisDayParams& p, S td,
                                         bool& sen_u, bool& sen_l) {
  S* se = (S*)p.seasonal;
  S& ax = se[v.i];
  S& su = se[v.dp + v.i];
  S& sl = se[2 * v.dp + v.i];
  const S tc = td - S(273.16);
  const S lat = S(p.lat_rad[v.i]);
  const S dl = daylength_hours(lat, S(p.jday));
  const bool dec = dl < daylength_hours(lat, S(p.jday - 1));
  ax = ax + (tc - ax) / S(p.acclim_tau);
  su = (dec && dl < S(p.sen_pstart_u)) ? su + mx(S(p.sen_tb_u) - tc, S(0)) : S(0);
  sl = (dec && dl < S(p.sen_pstart_l)) ? sl + mx(S(p.sen_tb_l) - tc, S(0)) : S(0);
  sen_u = p.sen_photoperiod_u && su >= S(p.sen_ycrit_u);
  sen_l = p.sen_photoperiod_l && sl >= S(p.sen_ycrit_l);
}

template <class S> IBIS_HD IBIS_INLINE void day_begin_cell(const DayView<S>& v, const IbisDayParams& p) {
  const S tmelt = S(p.tmelt), eps = S(p.epsilon);
  const S ddfac = S(1.0 / p.ddays);
  const S td = v.td();
  const S a10td = v.d(IBIS_DAY_A10TD);
  bool sen_u = false, sen_l = false;
  if (p.seasonal) seasonal_update(v, p, td, sen_u, sen_l);
  // cold phenology, upper canopy
  const S tthr_u = mx(mx(S(0) + tmelt, v.d(IBIS_DAY_TC) + S(p.tthresh_u_offset) + tmelt),
                      S(p.tthresh_u_abs) + tmelt);
  const S gthr_u = S(p.gthresh_u_c) + tmelt;
  const S agddu = (a10td < gthr_u || sen_u) ? S(0) : v.d(IBIS_DAY_AGDDU) + td - gthr_u;
  const S tempu = (a10td < tthr_u || sen_u)
                      ? mx(S(0), v.d(IBIS_DAY_TEMPU) - ddfac)
                      : mn(S(1), mx(S(0), agddu - S(p.gdd_thresh_u)) / S(p.gdd_ramp_u));
  // lower canopy
  const S tthr_l = S(p.tthresh_l_c) + tmelt;
  const S gthr_l = S(p.gthresh_l_c) + tmelt;
  const S agddl = (a10td < gthr_l || sen_l) ? S(0) : v.d(IBIS_DAY_AGDDL) + td - gthr_l;
  const S templ = (a10td < tthr_l || sen_l)
                      ? mx(S(0), v.d(IBIS_DAY_TEMPL) - ddfac)
                      : mn(S(1), mx(S(0), agddl - S(p.gdd_thresh_l)) / S(p.gdd_ramp_l));
  // drought phenology
  const S su = v.s(IBIS_P_SOIL_STRESSTU), sl = v.s(IBIS_P_SOIL_STRESSTL);
  const S dropu = drop_update(v.d(IBIS_DAY_DROPU), v.d(IBIS_DAY_A10ANCUB), su, p, ddfac);
  const S dropls = drop_update(v.d(IBIS_DAY_DROPLS), v.d(IBIS_DAY_A10ANCLS), sl, p, ddfac);
  const S dropl4 = drop_update(v.d(IBIS_DAY_DROPL4), v.d(IBIS_DAY_A10ANCL4), sl, p, ddfac);
  const S dropl3 = drop_update(v.d(IBIS_DAY_DROPL3), v.d(IBIS_DAY_A10ANCL3), sl, p, ddfac);
  // displayed leaf area and plant-type fractions
  S pl[IBIS_NPFT];
#pragma unroll
  for (int j = 0; j < IBIS_NPFT; ++j) pl[j] = v.d(IBIS_DAY_PLAI + j);
  const S avglaiu = pl[0] + pl[1] * dropu + pl[2] + pl[3] + pl[4] * tempu + pl[5] +
                    pl[6] * tempu + pl[7] * tempu;
  const S su_ = mx(avglaiu, eps);
  S fr[IBIS_NPFT];
  fr[0] = pl[0] / su_;
  fr[1] = pl[1] * dropu / su_;
  fr[2] = pl[2] / su_;
  fr[3] = pl[3] / su_;
  fr[4] = pl[4] * tempu / su_;
  fr[5] = pl[5] / su_;
  fr[6] = pl[6] * tempu / su_;
  fr[7] = pl[7] * tempu / su_;
  const S m9 = mn(templ, dropls), m10 = mn(templ, dropl4), m11 = mn(templ, dropl3);
  const S avglail = pl[8] + pl[9] * m9 + pl[10] * m10 + pl[11] * m11;
  const S sl_ = mx(avglail, eps);
  fr[8] = pl[8] / sl_;
  fr[9] = pl[9] * m9 / sl_;
  fr[10] = pl[10] * m10 / sl_;
  fr[11] = pl[11] * m11 / sl_;
  const S lai_l = mn(avglail / mx(v.s(IBIS_P_VEGETATION_FL), eps), S(12));
  const S lai_u = mn(avglaiu / mx(v.s(IBIS_P_VEGETATION_FU), eps), S(12));
  const S ztop_l = mn(mx(S(0.25), lai_l * S(0.25)), v.s(IBIS_P_VEGETATION_ZBOT + 1) - S(0.5));
  // store
  v.d(IBIS_DAY_TEMPU) = tempu;
  v.d(IBIS_DAY_TEMPL) = templ;
  v.d(IBIS_DAY_AGDDU) = agddu;
  v.d(IBIS_DAY_AGDDL) = agddl;
  v.d(IBIS_DAY_DROPU) = dropu;
  v.d(IBIS_DAY_DROPLS) = dropls;
  v.d(IBIS_DAY_DROPL3) = dropl3;
  v.d(IBIS_DAY_DROPL4) = dropl4;
#pragma unroll
  for (int j = 0; j < IBIS_NPFT; ++j) v.s(IBIS_P_VEGETATION_FRAC + j) = fr[j];
  v.s(IBIS_P_VEGETATION_LAI) = lai_l;
  v.s(IBIS_P_VEGETATION_LAI + 1) = lai_u;
  v.s(IBIS_P_VEGETATION_ZBOT) = S(0.05);
  v.s(IBIS_P_VEGETATION_ZTOP) = ztop_l;
  // the day's sums start from zero; the 10-day means stay in the day arena
#pragma unroll
  for (int k = 0; k < 7; ++k) v.reset(kDailySums[k]);
}

// ---- day end -------------------------------------------------------------------------------
// Lloyd-Taylor temperature factor and WFPS moisture factor (loops_gswp3.py:506-520).
template <class S> IBIS_HD IBIS_INLINE S lt_factor(S t, S tconst) {
  const S btemp = S(288.16), tzero = S(227.13);
  const S safe = mx(t, tzero + S(10.0));
  return mn(r_exp(tconst * (S(1.0) / (btemp - tzero) - S(1.0) / (safe - tzero))), S(10.0));
}
template <class S> IBIS_HD IBIS_INLINE S wfps_factor(S w, S wi, S wshift) {
  const S wfps = (S(1.0) - wi) * w * S(100.0) - wshift;
  const S wet = S(0.000371) * wfps * wfps - S(0.0748) * wfps + S(4.13);
  const S dry = r_exp((wfps - S(60.0)) * (wfps - S(60.0)) / S(-800.0));
  return wfps >= S(60.0) ? wet : dry;
}

// N mineralised (first) or immobilised (second) by decomposition of a pool
// (physics/biogeochem.py:181-186).
template <class S> struct NFlux { S min, imm; };
template <class S> IBIS_HD IBIS_INLINE NFlux<S> nmin_imm(S y, S cfrom, S cto, S out) {
  const S net = (S(1.0) / cto - y / cfrom) * out;
  const bool imm = y / cfrom > S(1.0) / cto;
  return {imm ? S(0) : net, imm ? net : S(0)};
}

// One day of soil C/N (physics/biogeochem.py:23-278); writes the carbon planes, returns
// tco2mic and tnmin and advances storedn / yrleach / ynleach.
template <class S>
IBIS_HD IBIS_INLINE void soilbgc_cell(const DayView<S>& v, const IbisDayParams& b, S& tco2mic,
                                      S& tnmin) {
  const S totts = S(1.0 / b.ndaypy);
  S cnr[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cnr[k] = S(b.cnr[k]);
  // litter splits (Verberne): parameters only, formed in FP64 (biogeochem.py:62-77)
  auto split = [&](double cn, double cn_struct, S& fl, S& fs, S& fm) {
    const double l = b.fmax * (cn * cn) / (b.rconst + cn * cn);
    const double st = (1.0 / cn - l / b.cnr[4] - (1.0 - l) / b.cnr[6]) /
                      (1.0 / cn_struct - 1.0 / b.cnr[6]);
    fl = S(l);
    fs = S(st);
    fm = S(1.0 - l - st);
  };
  S fll, fls, flm, frl, frs, frm, fwl, fws, fwm;
  split(b.cnleaf, b.cnr[5], fll, fls, flm);
  split(b.cnroot, b.cnr[5], frl, frs, frm);
  split(b.cnwood, b.cnr[7], fwl, fws, fwm);
  const S h0 = S(b.hsoi[0]), h1 = S(b.hsoi[1]), h2 = S(b.hsoi[2]), h3 = S(b.hsoi[3]);
  const S rdepth = S(1.0) / (h0 + h1 + h2 + h3);
  const S texfact = (S(0.01) * v.s(IBIS_P_SOIL_SAND) * h0 + S(0.01) * v.s(IBIS_P_SOIL_SAND + 1) * h1 +
                     S(0.01) * v.s(IBIS_P_SOIL_SAND + 2) * h2 +
                     S(0.01) * v.s(IBIS_P_SOIL_SAND + 3) * h3) * rdepth;
  const S fbpom = S(0.50);
  const S falll = v.c(IBIS_C_FALLL), fallr = v.c(IBIS_C_FALLR), fallw = v.c(IBIS_C_FALLW);
  const S totcin = (falll + fallr + fallw) * totts;
  const S slon = v.c(IBIS_C_CSOISLON), slop = v.c(IBIS_C_CSOISLOP), pas = v.c(IBIS_C_CSOIPAS);
  const S cmic = v.c(IBIS_C_TOTCMIC);
  const S totcbegin = v.c(IBIS_C_CLITLM) + v.c(IBIS_C_CLITLS) + v.c(IBIS_C_CLITRM) +
                      v.c(IBIS_C_CLITRS) + v.c(IBIS_C_CLITWM) + v.c(IBIS_C_CLITWS) + slop + slon +
                      pas + cmic + v.c(IBIS_C_CLITLL) + v.c(IBIS_C_CLITRL) + v.c(IBIS_C_CLITWL);
  const S cmicmx = S(b.fbsom) * totcbegin;
  const S cmicp = mn(cmic, cmicmx);
  const S cmicn = mx(S(0.0), cmic - cmicmx);
  const S lm = v.c(IBIS_C_CLITLM) + flm * falll * totts;
  const S ls = v.c(IBIS_C_CLITLS) + fls * falll * totts;
  const S ll = v.c(IBIS_C_CLITLL) + fll * falll * totts;
  const S rm = v.c(IBIS_C_CLITRM) + frm * fallr * totts;
  const S rs = v.c(IBIS_C_CLITRS) + frs * fallr * totts;
  const S rl = v.c(IBIS_C_CLITRL) + frl * fallr * totts;
  const S wm = v.c(IBIS_C_CLITWM) + fwm * fallw * totts;
  const S ws = v.c(IBIS_C_CLITWS) + fws * fallw * totts;
  const S wl = v.c(IBIS_C_CLITWL) + fwl * fallw * totts;
  const S dl = v.c(IBIS_C_DECOMPL), ds = v.c(IBIS_C_DECOMPS);
  const S o_rs = mn(ds * S(b.krs) * rs, rs);
  const S o_ws = mn(dl * S(b.kws) * ws, ws);
  const S o_ls = mn(dl * S(b.kls) * ls, ls);
  const S o_lm = mn(dl * S(b.klm) * lm, lm);
  const S o_rm = mn(ds * S(b.krm) * rm, rm);
  const S o_wm = mn(dl * S(b.kwm) * wm, wm);
  const S o_nb = mn(ds * S(b.knb) * slon, slon);
  const S o_pb = mn(ds * S(b.kpb) * slop, slop);
  const S o_sb = mn(ds * S(b.ksb) * pas, pas);
  const S o_bp = mn(S(b.kbp) * cmicp, cmicp) * S(b.effac);
  const S o_bn = mn(S(b.kbn) * cmicn, cmicn) * S(b.effac);
  const S dbdt = o_rs * S(b.yrs) + o_ws * S(b.yws) + o_ls * S(b.yls) + o_lm * S(b.ylm) +
                 o_rm * S(b.yrm) + o_wm * S(b.ywm) + o_nb * S(b.ynb) + o_pb * S(b.ypb) +
                 o_sb * S(b.ysb) - o_bp - o_bn;
  const S o_ll = mn(dl * S(b.kll) * ll, ll);
  const S o_rl = mn(ds * S(b.krl) * rl, rl);
  const S o_wl = mn(dl * S(b.kwl) * wl, wl);
  const S o_ns = mn(ds * S(b.kns) * slon, slon);
  const S lig = S(b.lig_frac) * (o_ll * S(b.yll) + o_rl * S(b.yrl) + o_wl * S(b.ywl));
  S dcndt = lig + (S(1.0) - fbpom) * (S(b.ybn) * o_bn + S(b.ybp) * o_bp) - o_nb - o_ns;
  const S o_ps = mn(ds * S(b.kps) * slop, slop);
  S dcpdt = lig + fbpom * (S(b.ybn) * o_bn + S(b.ybp) * o_bp) - o_pb - o_ps;
  const S fleach = S(0.30e-03 / 18.0) * (S(0.01) + S(0.04) * texfact);
  const S dcsdt = (S(b.yns) * o_ns + S(b.yps) * o_ps) - o_sb - fleach * pas;
  const S cleach = fleach * (pas + slop + slon);
  v.d(IBIS_DAY_YNLEACH) = v.d(IBIS_DAY_YNLEACH) + fleach * pas / cnr[1] +
                          fleach * slop / cnr[2] + fleach * slon / cnr[3];
  dcndt = dcndt - fleach * slon;
  dcpdt = dcpdt - fleach * slop;
  v.d(IBIS_DAY_YRLEACH) = v.d(IBIS_DAY_YRLEACH) + cleach;
  const NFlux<S> n[9] = {nmin_imm(S(b.yrs), cnr[0], cnr[5], o_rs),
                         nmin_imm(S(b.yls), cnr[0], cnr[5], o_ls),
                         nmin_imm(S(b.yws), cnr[0], cnr[7], o_ws),
                         nmin_imm(S(b.ywm), cnr[0], cnr[7], o_wm),
                         nmin_imm(S(b.ylm), cnr[0], cnr[6], o_lm),
                         nmin_imm(S(b.yrm), cnr[0], cnr[6], o_rm),
                         nmin_imm(S(b.ynb), cnr[0], cnr[3], o_nb),
                         nmin_imm(S(b.ypb), cnr[0], cnr[2], o_pb),
                         nmin_imm(S(b.ysb), cnr[0], cnr[1], o_sb)};
  S totimm = n[0].imm, totmin = n[0].min;
#pragma unroll
  for (int k = 1; k < 9; ++k) {
    totimm = totimm + n[k].imm;
    totmin = totmin + n[k].min;
  }
  const S lf = S(b.lig_frac);
  const S nrelps = o_ps * (S(1.0) / cnr[2] - S(1.0) / cnr[1]);
  const S nrelns = o_ns * (S(1.0) / cnr[3] - S(1.0) / cnr[1]);
  const S nrelbn = (S(1.0) - fbpom) * o_bn * (S(1.0) / cnr[0] - S(1.0) / cnr[3]) +
                   (S(1.0) - fbpom) * o_bp * (S(1.0) / cnr[0] - S(1.0) / cnr[3]);
  const S nrelbp = fbpom * o_bp * (S(1.0) / cnr[0] - S(1.0) / cnr[2]) +
                   fbpom * o_bn * (S(1.0) / cnr[0] - S(1.0) / cnr[2]);
  const S nrelll = lf * o_ll * (S(1.0) / cnr[4] - S(1.0) / cnr[2]) +
                   lf * o_ll * (S(1.0) / cnr[4] - S(1.0) / cnr[3]);
  const S nrelrl = lf * o_rl * (S(1.0) / cnr[4] - S(1.0) / cnr[2]) +
                   lf * o_rl * (S(1.0) / cnr[4] - S(1.0) / cnr[3]);
  const S nrelwl = lf * o_wl * (S(1.0) / cnr[4] - S(1.0) / cnr[2]) +
                   lf * o_wl * (S(1.0) / cnr[4] - S(1.0) / cnr[3]);
  const S totnrel = nrelps + nrelns + nrelbn + nrelbp + nrelll + nrelrl + nrelwl;
  const S netmin = totmin + totimm + totnrel;
  tnmin = mx(S(0.0), netmin) / S(86400.0 * 0.014);
  const S n_cmic = mx(S(0.0), cmic + dbdt);
  const S n_slon = mx(S(0.0), slon + dcndt);
  const S n_slop = mx(S(0.0), slop + dcpdt);
  const S n_pas = mx(S(0.0), pas + dcsdt);
  const S n_lm = mx(S(0.0), lm - o_lm), n_ls = mx(S(0.0), ls - o_ls), n_ll = mx(S(0.0), ll - o_ll);
  const S n_rm = mx(S(0.0), rm - o_rm), n_rs = mx(S(0.0), rs - o_rs), n_rl = mx(S(0.0), rl - o_rl);
  const S n_wm = mx(S(0.0), wm - o_wm), n_ws = mx(S(0.0), ws - o_ws), n_wl = mx(S(0.0), wl - o_wl);
  const S totlit = n_lm + n_ls + n_ll + n_rm + n_rs + n_rl + n_wm + n_ws + n_wl;
  const S totalit = n_lm + n_ls + n_wm + n_ll + n_ws + n_wl;
  const S totrlit = n_rm + n_rs + n_rl;
  const S totcsoi = n_pas + n_slop + n_cmic + n_slon;
  const S totcend = totlit + totcsoi;
  tco2mic = (totcbegin + totcin - totcend - cleach) / S(86400.0 * 0.012);
  v.d(IBIS_DAY_STOREDN) = v.d(IBIS_DAY_STOREDN) + tnmin;
  v.c(IBIS_C_CLITLM) = n_lm; v.c(IBIS_C_CLITLS) = n_ls; v.c(IBIS_C_CLITLL) = n_ll;
  v.c(IBIS_C_CLITRM) = n_rm; v.c(IBIS_C_CLITRS) = n_rs; v.c(IBIS_C_CLITRL) = n_rl;
  v.c(IBIS_C_CLITWM) = n_wm; v.c(IBIS_C_CLITWS) = n_ws; v.c(IBIS_C_CLITWL) = n_wl;
  v.c(IBIS_C_CSOIPAS) = n_pas; v.c(IBIS_C_CSOISLON) = n_slon; v.c(IBIS_C_CSOISLOP) = n_slop;
  v.c(IBIS_C_TOTCMIC) = n_cmic;
  v.c(IBIS_C_TOTLIT) = totlit; v.c(IBIS_C_TOTALIT) = totalit; v.c(IBIS_C_TOTRLIT) = totrlit;
  v.c(IBIS_C_TOTCSOI) = totcsoi; v.c(IBIS_C_TOTFALL) = falll + fallr + fallw;
}

template <class S>
IBIS_HD IBIS_INLINE void day_end_cell(const DayView<S>& v, S* vegx, int64_t vpitch,
                                      const IbisDayParams& p) {
  // 10-day means from the day's sums (loops_gswp3.py:470-477)
  const S z = r_exp(S(-1.0) / S(10.0));
  const S al = S(1.0) - z;
  const S day_sec = S(p.day_sec);
  v.d(IBIS_DAY_A10TD) = z * v.d(IBIS_DAY_A10TD) + al * v.td();
  v.d(IBIS_DAY_A10ANCUB) = z * v.d(IBIS_DAY_A10ANCUB) + al * (v.sum(IBIS_ACC_A10ANCUB) / day_sec);
  v.d(IBIS_DAY_A10ANCLS) = z * v.d(IBIS_DAY_A10ANCLS) + al * (v.sum(IBIS_ACC_A10ANCLS) / day_sec);
  v.d(IBIS_DAY_A10ANCL3) = z * v.d(IBIS_DAY_A10ANCL3) + al * (v.sum(IBIS_ACC_A10ANCL3) / day_sec);
  v.d(IBIS_DAY_A10ANCL4) = z * v.d(IBIS_DAY_A10ANCL4) + al * (v.sum(IBIS_ACC_A10ANCL4) / day_sec);
  v.d(IBIS_DAY_AYPRCP) = v.d(IBIS_DAY_AYPRCP) + v.sum(IBIS_ACC_ADRAIN) + v.sum(IBIS_ACC_ADSNOW);
  v.d(IBIS_DAY_AYAET) = v.d(IBIS_DAY_AYAET) + v.sum(IBIS_ACC_ADAET);
  // decomposition climate factors (loops_gswp3.py:494-553)
  const S tconst = S(p.lt_e0), wshift = S(p.wfps_opt) - S(60.0);
  const S fl_top = lt_factor(v.s(IBIS_P_SOIL_TSOI), tconst) *
                   wfps_factor(v.s(IBIS_P_SOIL_WSOI), v.s(IBIS_P_SOIL_WISOI), wshift);
  v.c(IBIS_C_DECOMPL) = clip(fl_top * S(p.k_scale), S(0.001), S(10.0));
  const S h0 = S(p.hsoi[0]), h1 = S(p.hsoi[1]), h2 = S(p.hsoi[2]), h3 = S(p.hsoi[3]);
  const S h4 = h0 + h1 + h2 + h3;
  auto avg = [&](int pl) {
    return (v.s(pl) * h0 + v.s(pl + 1) * h1 + v.s(pl + 2) * h2 + v.s(pl + 3) * h3) / h4;
  };
  const S fs = lt_factor(avg(IBIS_P_SOIL_TSOI), tconst) *
               wfps_factor(avg(IBIS_P_SOIL_WSOI), avg(IBIS_P_SOIL_WISOI), wshift);
  v.c(IBIS_C_DECOMPS) = clip(fs * S(p.k_scale), S(0.001), S(10.0));
  // soil biogeochemistry, nsoilcall times (loops_gswp3.py:564-575)
  S tco2mic = S(0), tnmin = S(0);
  for (int k = 0; k < p.nsoilcall; ++k) soilbgc_cell(v, p, tco2mic, tnmin);
  v.d(IBIS_DAY_TNMIN) = tnmin;
  vegx[(int64_t)IBIS_VEGX_TCO2MIC * vpitch + v.i] = tco2mic;
  // annual heterotrophic and root respiration (loops_gswp3.py:577-585)
  v.d(IBIS_DAY_AYCO2MIC) = v.d(IBIS_DAY_AYCO2MIC) + tco2mic * S(12.0e-3 * 86400.0);
  v.d(IBIS_DAY_AYCO2ROOT) = v.d(IBIS_DAY_AYCO2ROOT) + v.d(IBIS_DAY_TCO2ROOT) * S(12.0e-3 * 86400.0);
  // soil-water stress of the tree-establishment gate, growing-season days only
  // (loops_gswp3.py:587-599)
  if (p.stress) {
    S* sa = (S*)p.stress;
    const S grow = v.td() > S(p.tmelt + p.stress_accum_tmin) ? S(1) : S(0);
    sa[v.i] = sa[v.i] + v.s(IBIS_P_SOIL_STRESSTU) * grow;
    sa[v.dp + v.i] = sa[v.dp + v.i] + v.s(IBIS_P_SOIL_STRESSTL) * grow;
    sa[2 * v.dp + v.i] = sa[2 * v.dp + v.i] + grow;
  }
}

}  // namespace ibis
