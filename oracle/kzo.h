/* kzo.h — the CPU restatement ("oracle") of koral-zig's per-step hot path.
 *
 * TEST INFRASTRUCTURE.  This directory is the correctness oracle and the
 * CPU-baseline path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (koral-zig_amd/) never does.
 *
 * Every function restates one reference function op-for-op (same operand
 * order, same association, same branch structure) so that the HIP path can
 * be checked against it to the last bit; each carries the file:line it
 * follows.  Transcendentals come from kz_math.h, shared with the kernels.
 *
 * Parity pins (tests/test_oracle_golden.py): every function below is checked
 * against the reference's committed C-oracle golden records under
 * tests/golden/ (recon, wavespeed, fluxprime, p2u, u2p_solver, floors, Tij,
 * rad_rij, rad_u2prad, rad_wavespeeds, rad_floors, rad_thermo, rad_opac,
 * rad_gi, rad_implicit, ct) and against closed forms (LAXF F*(U,U)=F(U),
 * invert4 vs numpy, p2u/u2p round trips).
 */
#ifndef KZO_H
#define KZO_H

#include <stdint.h>
#include "../include/kz_types.h"

#define KZM_QUAL static inline
#include "../koral-zig_amd/csrc/kz_math.h"

#define NV KZ_NV
#define RHO KZ_RHO
#define UU KZ_UU
#define VX KZ_VX
#define VY KZ_VY
#define VZ KZ_VZ
#define ENTR KZ_ENTR
#define B1 KZ_B1
#define B2 KZ_B2
#define B3 KZ_B3
#define EE KZ_EE
#define FX KZ_FX
#define FY KZ_FY
#define FZ KZ_FZ

/* point geometry view (koral/geometry.zig:29-49) */
typedef struct {
  double g[4][4];
  double G[4][4];
  double gdet, alpha, gttpert;
  double x1; /* internal radial coordinate xxvec[1] (horizon_x1 test) */
} kzo_geo;

/* relele.zig velocity representations */
enum { VEL4 = 1, VEL3 = 2, VELR = 3 };

/* error codes mirror relele.Error */
#define KZO_OK 0
#define KZO_ESPACELIKE KZ_E_SPACELIKE
#define KZO_EVELCONV KZ_E_VELCONV
#define KZO_ENANFLUX KZ_E_NAN_IN_FLUX

typedef struct {
  double gas_abs, rad_abs, gas_num, rad_num, gas_ross, rad_ross, tot_emissivity;
} kzo_opac;

/* radforce.RadStateOf (koral/physics/radforce.zig:156-178) */
typedef struct {
  double rho, uint, tgas, te, ti, sgas, ne;
  double ucon[4], ucov[4];
  double bsq;
  double rij[4][4];
  double ehat, tradbb, trad, kappaes, kappa;
  kzo_opac opac;
} kzo_radstate;

enum { OPAC_FULL = 0, OPAC_RESIDUAL = 1, OPAC_CHI = 2 };
enum { WP_RAD = 1, WP_MHD = 2 };
enum { WE_ENERGY = 0, WE_ENTROPY = 1 };
enum { WF_LAB = 0, WF_FF = 1 };

/* the reductions of one scalars.dat row (io/scalars.zig), see kzo_sim.c */
typedef struct kzo_sim kzo_sim;
int kzo_sim_scalars(kzo_sim *s, long ix_h, long ix_l, long ix_s, const double *dth, const double *dph, double out[6],
                    uint64_t cnt[2]);

/* threads of the oracle sim's cell-parallel passes (results independent of it) */
void kzo_set_threads(int n);
int kzo_get_threads(void);

/* ---- metric.zig:20-32 ---- */
double kzo_r_horizon_bl(double a);
double kzo_r_isco_bl(double a);

/* ---- geometry ---- */
void kzo_geo_from_kz(const kz_geom *k, kzo_geo *o);
void kzo_geo_from_blocks(const double *g20, const double *G20, double x1, kzo_geo *o);

/* ---- relele.zig ---- */
void kzo_raise(const double a[4], const kzo_geo *gm, double out[4]);
void kzo_lower(const double a[4], const kzo_geo *gm, double out[4]);
double kzo_dot(const double a[4], const double b[4]);
void kzo_lower_second(const double t[4][4], const kzo_geo *gm, double out[4][4]);
void kzo_lower_second_row(const double t[4][4], const kzo_geo *gm, int row, double out[4]);
double kzo_alpha_gamma(const double u[4], const kzo_geo *gm);
void kzo_velr_to_vel4(const double uin[4], const kzo_geo *gm, double out[4]);
int kzo_convert(const double uin[4], int from, int to, const kzo_geo *gm, int trust_ut, double out[4]);
void kzo_ucon_ucov_from_prims(const double *pp, const kzo_geo *gm, double ucon[4], double ucov[4]);

/* ---- bfield.zig / hydro.zig / p2u.zig ---- */
void kzo_bcon_bcov_bsq(const double B[3], const double ucon[4], const double ucov[4],
                       const kzo_geo *gm, double bcon[4], double bcov[4], double *bsq);
double kzo_s_from_u(double rho, double u, double gam);
double kzo_calc_utp1(const double vcon[4], const kzo_geo *gm);
void kzo_p2u_mhd(const double *pp, double *uu, const kzo_geo *gm, double gam);
void kzo_p2u_rad(const double *pp, double *uu, const kzo_geo *gm);
void kzo_p2u(const double *pp, double *uu, const kzo_geo *gm, double gam);
int kzo_calc_tij(const double *pp, const kzo_geo *gm, double gam, double t[4][4]);

/* ---- thermo.zig / radiation.zig ---- */
double kzo_t_from_urho(const kz_opac_params *c, double uint, double rho, double gam);
double kzo_lte_t_from_e(const kz_opac_params *c, double e);
void kzo_calc_rij(const double *pp, const kzo_geo *gm, double rij[4][4]);
double kzo_calc_ff_rtt(const double *pp, const kzo_geo *gm, double ucon_out[4]);
void kzo_lr_core(const double ucon[4], const kzo_geo *gm, const double ws2[3], const int active[3],
                 double aret[6]);
void kzo_gas_wavespeeds_lr(const double *pp, const kzo_geo *gm, double gam, const int active[3],
                           double a[6]);
void kzo_calc_rad_wavespeeds(const double *pp, const kzo_geo *gm, const double tautot[3],
                             const int active[3], double aval[12]);

/* ---- opacities.zig / radforce.zig ---- */
void kzo_calc_opacities(const kz_opac_params *c, double rho, double tgas, double te, double trad,
                        double tradbb, double ne, double bsq, int level, kzo_opac *o);
double kzo_kappa_es_puffy(const kz_opac_params *c, double rho, double trad);
void kzo_fill_radstate(const double *pp, const kzo_geo *gm, double gam, const kz_opac_params *c,
                       int full, kzo_radstate *st);
void kzo_calc_gi(const kzo_radstate *st, const double *pp, const kzo_geo *gm,
                 const kz_opac_params *c, int full, double gi_ff[4], double gi_lab[4]);
double kzo_compton_coeff(const kz_opac_params *c, const kzo_radstate *st);
double kzo_calc_kappaes_standalone(const double *pp, double gam, const kz_opac_params *c);
double kzo_calc_chi_slim(const double *pp, const kzo_geo *gm, double gam, const kz_opac_params *c);

/* ---- solve/invert.zig ---- */
int kzo_u2p_solver_w(const double *uu, double *pp, const kzo_geo *gm, int etype, double gam);
kz_u2p_result kzo_u2p_mhd(const double *uu, double *pp, const kzo_geo *gm, double gam,
                          const kz_floor_params *f);
int kzo_check_floors_mhd(double *pp, const kzo_geo *gm, double gam, const kz_floor_params *f,
                         uint64_t *n_guard);

/* ---- solve/invert_rad.zig ---- */
kz_u2prad_result kzo_u2p_rad(const double *uu, double *pp, const kzo_geo *gm, const kz_rad_params *rp);
int kzo_check_floors_rad(double *pp, const kzo_geo *gm, const kz_rad_params *rp);

/* ---- solve/implicit.zig ---- */
void kzo_invert4(const double a[4][4], double inv[4][4]);
void kzo_invert4f(const float a[4][4], float inv[4][4]);
kz_implicit_result kzo_solve_implicit_lab(double *uu, double *pp, const kzo_geo *gm, double dt,
                                          double gam, const kz_rad_params *rp,
                                          const kz_opac_params *op, const kz_implicit_params *ip);

/* ---- fv/recon.zig, fv/laxf.zig, physics/flux.zig ---- */
void kzo_reconstruct(int order, double theta, const double u[5], const double dx[5],
                     double *left, double *right);
double kzo_laxf(double fl, double fr, double ul, double ur, double ag);
int kzo_f_flux_prime(const double *pp, int idim, const kzo_geo *gm, double gam, double *ff);

#endif /* KZO_H */
