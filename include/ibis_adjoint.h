/*
 * C ABI of the parameter-estimation step: the hourly step with a per-pair tunable parameter
 * vector, the hourly flux misfit against target fluxes, and its forward-mode derivative with
 * respect to the tunable parameters.
 *
 * Threads.  A launch covers ndir x npair threads; thread t = k * npair + p steps pair p
 * (0 <= p < npair) and, in tangent mode, carries the derivative along tunable component
 * dir0 + k.  Pair p runs site p % nsite: the site's latitude, slot types, forcing (the
 * benchmark generator of ibis_forcing_synthetic keyed by seed, site0 + site index and hour) and
 * target fluxes; its parameter vector is column p of `tune`.  Running variants of a site's
 * vector as extra pairs (npair = nsite x variants) evaluates finite differences in one launch.
 *
 * State.  Each thread owns its cell of the arena (pitch >= ndir * npair).  In value mode the
 * elements are float (IBIS_PREC_MIXED) or double (IBIS_PREC_F64); in tangent mode
 * (IBIS_FLAG_TANGENT) each element is a (value, tangent) pair of that type, interleaved, so
 * that the arena is one of 2-element structures.  The value part of a tangent run performs
 * the same operations as the value run.
 *
 * Loss.  After the step, the hour's fluxes of each thread are formed: GPP (mol CO2 m-2 s-1,
 * the driver's canopy gross assimilation, python-port/ibis/timestepping/loops_gswp3.py:239-252),
 * latent heat LE = hvap (ginvap + gsuvap + gtrans) (W m-2; canopy_exchange.py:794-798) and
 * sensible heat H = fsena (W m-2, the total to the air after cascad2).  With `flux` given they
 * are stored (float, IBIS_NFLUX planes of `pitch`); with `target` given (float, IBIS_NFLUX
 * planes of `nsite`), sum_f ((flux_f - target_f) / scale_f)^2 is added to loss[t] (double)
 * and, in tangent mode, theta_k dL/dtheta_k (the derivative along the seed theta_k e_k, k =
 * dir0 + t / npair) to loss[pitch + t].
 *
 * No counterpart in the reference's step (its gradients come from jax.grad over the scan).
 * Stream-ordered; allocates nothing.
 */
#ifndef IBIS_ADJOINT_H
#define IBIS_ADJOINT_H

#include <stdint.h>
#include "ibis_step.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tunable parameters, in order: alpha3 alpha4 theta3 theta4 beta3 beta4 coefm[5] coefb[5]
 * stressfac cleaf cstem cgrass chu chs chl wliqumax wsnoumax fmax bperm wpudmax bed_consoi
 * bed_csoi, then the 12 plant-type Vmax (python-port/ibis/physics/stomata.py:49-91, :84-88;
 * canopy_exchange.py:70-79; interception.py:33-45; timestepping/loops.py:92-94;
 * physics/bedrock.py:42-45). */
#define IBIS_NTUNE 42
#define IBIS_TUNE_VMAX 30
#define IBIS_NFLUX 3 /* GPP, LE, H */

/* Element type holds (value, tangent) and the tangent along dir0 + k is propagated. */
#define IBIS_FLAG_TANGENT 0x40

typedef struct IbisTuned {
  int64_t nsite, npair, ndir;  /* threads = ndir * npair */
  int32_t dir0;                /* tangent mode: first seeded component */
  int32_t jday0;               /* day of year of hour 0 */
  uint64_t seed;               /* forcing seed */
  int64_t hour;                /* hour index of this step (forcing) */
  double dtime, nytimes;       /* step length; VegetationState.nytimes before the step */
  double flux_scale[IBIS_NFLUX];
  int64_t site0;               /* forcing key of site 0: site i draws the forcing of cell
                                  site0 + i (a window of a larger seeded domain) */
} IbisTuned;

/* Workspace bytes for arenas of `pitch` threads under `flags` (precision | TANGENT). */
int64_t ibis_tuned_workspace_bytes(int64_t pitch, int flags);

/* One hour of every thread: forcing, the step with the pair's parameters, fluxes / misfit.
 * state: arena described above; slot_type: IBIS_NPFT int8 planes of nsite; lat_deg: nsite
 * doubles; tune: IBIS_NTUNE double planes of npair; row_pft: IBIS_TYPE_ROWS int32, the
 * plant-type position whose Vmax each type-table row takes (python-port/ibis/io/pft_types.py:
 * 170-189); base: every other parameter.  flux / target / loss may be NULL.  Device pointers
 * except base and run. */
int ibis_tuned_hour(const IbisArena* state, void* workspace, int64_t workspace_bytes,
                    const int8_t* slot_type, const double* lat_deg, const double* tune,
                    const int32_t* row_pft, const IbisParams* base, const IbisTuned* run,
                    const float* target, float* flux, double* loss, int flags,
                    void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* IBIS_ADJOINT_H */
