/* kz_cellops.h — batched per-point functions of the step on the GPU
 * (libkz_hip.so).
 *
 * Each entry point evaluates one per-cell (or per-face) function of the
 * reference over n independent records, one GPU thread per record, with the
 * device code the step kernels use.  Geometry arrives as kz_geom records
 * (koral/geometry.zig:29-49), states as 13 doubles per record in the slot
 * order of kz_types.h.  Buffers are host memory; every call is synchronous.
 * Return: KZ_OK, KZ_E_HIP on a runtime failure, KZ_E_INVALID_CONFIG for
 * arguments outside the reference's ranges (an inversion type, dimension or
 * function index out of range, a b^2/rho floor frame other than drift or
 * ZAMO).
 */
#ifndef KZ_CELLOPS_H
#define KZ_CELLOPS_H

#include <stdint.h>
#include "kz_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* p2u (koral/p2u.zig:135-148): primitives -> conserved.  The primitives'
 * velocities are VELR, whose conversion to u^mu cannot fail
 * (koral/relele.zig:360-367 convertBothVelrG), so p2u has no per-cell error
 * to report and the batch has no error output. */
int kz_p2u_batch(int64_t n, const double *pp, const kz_geom *g, double gam, double *uu);

/* u2pMhd (koral/solve/invert.zig:422-472): pp in = guess, out = result. */
int kz_u2p_mhd_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, double gam, const kz_floor_params *f,
                     kz_u2p_result *r);

/* checkFloorsMhd (koral/solve/invert.zig:474-715), drift-frame branch;
 * *nguard = number of records whose floor pass took the guard ladder. */
int kz_check_floors_mhd_batch(int64_t n, double *pp, const kz_geom *g, double gam, const kz_floor_params *f,
                              int32_t *ret, uint64_t *nguard);

/* u2pRad (koral/solve/invert_rad.zig:189-264). */
int kz_u2p_rad_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, const kz_rad_params *rp,
                     kz_u2prad_result *r);

/* checkFloorsRad (koral/solve/invert_rad.zig:266-330). */
int kz_check_floors_rad_batch(int64_t n, double *pp, const kz_geom *g, const kz_rad_params *rp, int32_t *ret);

/* solveImplicitLab (koral/solve/implicit.zig:927-1009): the 6-rung Newton
 * ladder for the radiation-gas four-force over dt[i]; uu/pp updated in place
 * when r[i].ok. */
int kz_solve_implicit_lab_batch(int64_t n, double *uu, double *pp, const kz_geom *g, const double *dt, double gam,
                                const kz_rad_params *rp, const kz_opac_params *op, const kz_implicit_params *ip,
                                kz_implicit_result *r);

/* reconstruct (koral/fv/recon.zig:47-53): order 0 donor cell, 1 linear
 * minmod with theta (:77-91), 2 PPM (:99-195); 5-point stencil values and
 * cell sizes -> left/right face values of the centre cell. */
int kz_reconstruct_batch(int64_t n, const int32_t *order, const double *theta, const double *u5, const double *dx5,
                         double *l, double *r);

/* u2pSolverW (koral/solve/invert.zig:235-404): the W Newton alone,
 * etype 0 hot (energy) / 1 entropy; pp in = guess, out = result. */
int kz_u2p_solver_w_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, int32_t etype, double gam,
                          int32_t *ret);

/* fFluxPrime (koral/physics/flux.zig:40-134); err[i] = KZ_E_NAN_IN_FLUX
 * when a component is not finite. idim in 0..2. */
int kz_flux_prime_batch(int64_t n, const double *pp, const int32_t *idim, const kz_geom *g, double gam, double *ff,
                        int32_t *err);

/* gasWavespeedsLr (koral/physics/wavespeeds.zig:98-132): a6 = l/r per dim;
 * act = the three active-dimension switches. */
int kz_gas_wavespeeds_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const int32_t *act,
                            double *a6);

/* calcRadWavespeeds (koral/physics/radiation.zig:177-211): a12 = the
 * optically thin l/r set then the tau-limited set. */
int kz_rad_wavespeeds_batch(int64_t n, const double *pp, const kz_geom *g, const double *tau3, const int32_t *act,
                            double *a12);

/* calcRijG + calcFfRtt (koral/physics/radiation.zig:79-124):
 * out21 = R^{ij}[16], R^t_t(fluid frame), u^mu[4]. */
int kz_rij_batch(int64_t n, const double *pp, const kz_geom *g, double *out21);

/* fillRadStateG / fillRadStateSlimG (koral/physics/radforce.zig:187-260) and
 * calcChiSlim (radforce.zig:562-590): out10 = Tgas Te Ehat Trad kappaes
 * kappa_gas_abs kappa_rad_abs kappa_rad_ross sgas chi_slim. */
int kz_radstate_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const kz_opac_params *op,
                      int32_t full, double *out10);

/* The transcendental and IEEE primitives the device code is built on, for
 * bit-level checks: fn 0 exp, 1 log, 2 pow(x, y), 3 cbrt, 4 log1p, 5 sqrt,
 * 6 x / y. */
int kz_math_batch(int64_t n, int32_t fn, const double *x, const double *y, double *out);

#ifdef __cplusplus
}
#endif

#endif /* KZ_CELLOPS_H */
